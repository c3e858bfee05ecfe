// solver.hip -- flexible GCR and the SAP+GCR driver.
//
// Replaces linsolv/fgcr.c:109-299 and sap/sap_gcr.c:120-254.  The algorithm
// and its precision structure are those of the reference: single-precision
// Krylov vectors phi (preconditioned directions) and chi (their images),
// modified Gram-Schmidt with double-precision accumulation, restart after nkv
// vectors with the solution update psi += sum c_l phi_l in double precision
// and the residual recomputed as rho = eta - D psi with the double-precision
// operator.  Stopping: rn <= res*rn0, stagnation rn < 100*FLT_EPSILON*rn_old,
// k = nkv-1 or nmx Krylov vectors in total.  All vectors are device fields;
// the host keeps only the small arrays a, b, c.
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "oqh_internal.h"

namespace oqh {
bool sap_prepare(int isolv, const char *where, bool quiet);
bool sap_mixed_mop(float mu, int isolv, int nmr, int ncy, const float2 *rho, float2 *phi, bool want_res);
void sap_mixed_residual(float2 *r);
bool sap_mop64(double mu, int isolv, int nmr, int ncy, const float2 *rho, float2 *phi, float2 *chi);
bool sap_multi_device(float mu, int isolv, int nmr, int ncy, int n, float2 *const *psi, float2 *const *rho);
}

using namespace oqh;

static int dummy_off = 0;

// device slots (and a pinned host mirror) for the coefficients of the fused
// Gram-Schmidt steps
static void *tmp_gs_coef(int n, float2 **host)
{
   static float2 *d = nullptr, *h = nullptr;
   static int cap = 0;
   if (n > cap)
   {
      dev_free(d);
      if (h) (void)hipHostFree(h);
      d = (float2 *)dev_alloc(sizeof(float2) * n);
      h = nullptr;
      if (hipHostMalloc((void **)&h, sizeof(float2) * n, 0) != hipSuccess) h = nullptr;
      cap = (d && h) ? n : 0;
   }
   *host = h;
   return cap ? d : nullptr;
}

static double get_res(int vol, int istop, spinor *rho)
{
   if (istop) return (double)unorm(vol, 3, rho);
   double rn = (double)norm_square(vol, 3, rho);
   return sqrt(rn);
}

extern "C" double fgcr(int vol, void (*Dop)(spinor_dble *s, spinor_dble *r),
                       void (*Mop)(int k, spinor *rho, spinor *phi, spinor *chi), spinor **ws,
                       spinor_dble **wsd, int nkv, int nmx, int istop, double res, spinor_dble *eta,
                       spinor_dble *psi, int *status)
{
   if (vol <= 0 || nkv < 1 || nmx < 1 || res <= DBL_EPSILON || istop < 0 || istop > 1)
   {
      fail("fgcr [fgcr.c]", "Parameters are out of range");
      return 0.0;
   }
   std::vector<complex> a((size_t)nkv * nkv), c(nkv);
   std::vector<float> b(nkv);
   spinor **phi = ws, *rho = ws[nkv], **chi = ws + nkv + 1;
   spinor_dble *wrk = wsd[0];
   const double PRECISION_LIMIT = (double)(100.0f * FLT_EPSILON);

   set_sd2zero(vol, 2, psi);
   assign_sd2s(vol, 2, eta, rho);
   double rn = get_res(vol, istop, rho), tol = res * rn, rn_old;
   status[0] = 0;

   // device path for the Gram-Schmidt steps: all Krylov vectors and rho are
   // full device fields and the stopping norm is the 2-norm
   Lib &L = lib();
   int npts = (int)((long)vol * L.g.nthread);
   bool fused = (istop == 0 && npts == L.g.V);
   for (int l = 0; fused && l < 2 * nkv + 1; l++)
   {
      int off = 0;
      fused = device_field(ws[l], F32, &off) && off == 0;
   }
   const int SB_ = nkv, SC_ = nkv + 1, SR_ = nkv + 2;
   float2 *dcoef = nullptr, *hcoef = nullptr;
   static float2 *d_gram = nullptr;
   static int gram_ld = 0;
   if (fused)
   {
      dcoef = (float2 *)tmp_gs_coef(nkv + 3, &hcoef);
      if (nkv > gram_ld)
      {
         dev_free(d_gram);
         gram_ld = nkv;
         d_gram = (float2 *)dev_alloc(sizeof(float2) * gram_ld * gram_ld);
         if (!d_gram) gram_ld = 0;
      }
      fused = dcoef != nullptr && d_gram != nullptr;
   }
   auto F = [](spinor *p) { return (float2 *)device_field(p, F32, &dummy_off); };
   std::vector<float2 *> qv(nkv);

   while (rn > tol)
   {
      int k;
      rn_old = rn;
      for (k = 0;; k++)
      {
         if (fused)
         {
            // gcr_step with the coefficients kept on the device: one fused
            // update-and-product pass per Gram-Schmidt term, one read-back
            Mop(k, rho, phi[k], chi[k]);
            float2 *ck = F(chi[k]);
            if (L.gcr_cgs)
            {
               for (int l = 0; l < k; l++) qv[l] = F(chi[l]);
               gs_classical(ck, qv.data(), k, F(rho), dcoef, SB_, SC_, SR_, npts, d_gram, gram_ld);
            }
            else if (k > 0)
            {
               gs_step<GS_DOT>(ck, nullptr, F(chi[0]), dcoef, 0, 0, npts);
               for (int l = 0; l < k; l++)
               {
                  if (l + 1 < k)
                     gs_step<GS_AXPY_DOT>(ck, F(chi[l]), F(chi[l + 1]), dcoef, l, l + 1, npts);
                  else
                     gs_step<GS_AXPY_NORM>(ck, F(chi[l]), nullptr, dcoef, l, SB_, npts);
               }
            }
            else
               gs_step<GS_NORM>(ck, nullptr, nullptr, dcoef, 0, SB_, npts);
            if (!L.gcr_cgs)
            {
               gs_step<GS_SCALE_DOT>(ck, nullptr, F(rho), dcoef, SB_, SC_, npts);
               gs_step<GS_AXPY_NORM>(F(rho), ck, nullptr, dcoef, SC_, SR_, npts);
            }
            HIP_OK(hipMemcpyAsync(hcoef, dcoef, sizeof(float2) * (nkv + 3), hipMemcpyDeviceToHost, L.stream));
            HIP_OK(hipStreamSynchronize(L.stream));
            for (int l = 0; l < k; l++)
            {
               a[(size_t)nkv * l + k].re = hcoef[l].x;
               a[(size_t)nkv * l + k].im = hcoef[l].y;
            }
            b[k] = hcoef[SB_].x;
            if (b[k] == 0.0f) fail("normalize [salg.c]", "Vector has vanishing norm");
            c[k].re = hcoef[SC_].x;
            c[k].im = hcoef[SC_].y;
            rn = sqrt((double)hcoef[SR_].y);
            status[0] += 1;
            if (rn <= tol || rn < PRECISION_LIMIT * rn_old || k == nkv - 1 || status[0] >= nmx) break;
            continue;
         }
         // gcr_step (fgcr.c:157-178)
         Mop(k, rho, phi[k], chi[k]);
         for (int l = 0; l < k; l++)
         {
            a[(size_t)nkv * l + k] = spinor_prod(vol, 3, chi[l], chi[k]);
            complex z;
            z.re = -a[(size_t)nkv * l + k].re;
            z.im = -a[(size_t)nkv * l + k].im;
            mulc_spinor_add(vol, 2, chi[k], chi[l], z);
         }
         b[k] = normalize(vol, 3, chi[k]);
         c[k] = spinor_prod(vol, 3, chi[k], rho);
         complex z;
         z.re = -c[k].re;
         z.im = -c[k].im;
         mulc_spinor_add(vol, 2, rho, chi[k], z);
         rn = get_res(vol, istop, rho);
         status[0] += 1;
         if (rn <= tol || rn < PRECISION_LIMIT * rn_old || k == nkv - 1 || status[0] >= nmx) break;
      }
      // update_psi (fgcr.c:181-212)
      for (int l = k; l >= 0; l--)
      {
         float zr = c[l].re, zi = c[l].im;
         for (int i = l + 1; i <= k; i++)
         {
            zr -= (a[(size_t)l * nkv + i].re * c[i].re - a[(size_t)l * nkv + i].im * c[i].im);
            zi -= (a[(size_t)l * nkv + i].re * c[i].im + a[(size_t)l * nkv + i].im * c[i].re);
         }
         float r = 1.0f / b[l];
         c[l].re = zr * r;
         c[l].im = zi * r;
      }
      set_s2zero(vol, 2, rho);
      for (int l = k; l >= 0; l--) mulc_spinor_add(vol, 2, rho, phi[l], c[l]);
      add_s2sd(vol, 2, rho, psi);
      Dop(psi, wrk);
      diff_sd2s(vol, 2, eta, wrk, rho);
      rn = get_res(vol, istop, rho);
      if (status[0] >= nmx && rn > tol)
      {
         status[0] = -1;
         return rn;
      }
   }
   return rn;
}

// ---------------------------------------------------------------------------
// sap_gcr
// ---------------------------------------------------------------------------

static double mud;
static bool nn_use = false;  // network correction in Mop for the current solve

static void Dop(spinor_dble *s, spinor_dble *r) { Dw_dble(mud, s, r); }

static void Mop(int k, spinor *rho, spinor *phi, spinor *chi)
{
   (void)k;
   Lib &L = lib();
   int vt = L.g.VT;
   int o0 = 0, o1 = 0, o2 = 0;
   float2 *r = (float2 *)device_field(rho, F32, &o0), *p = (float2 *)device_field(phi, F32, &o1),
          *c = (float2 *)device_field(chi, F32, &o2);
   const bool dev = r && p && c && !o0 && !o1 && !o2;
   if (L.sap_prec == 1 && dev)
   {
      // FP16 variant: the whole preconditioner on FP16 fields; with the
      // network, phi += NN(residual of the Schwarz cycles); chi = D phi with
      // the FP32 operator
      if (sap_mixed_mop((float)mud, L.sap_isolv, L.sap_nmr, L.sap_ncy, r, p, nn_use))
      {
         if (nn_use) sap_mixed_residual(c), nn_correct(c, p);
         Dw((float)mud, phi, chi);
         return;
      }
   }
   if (L.sap_prec == 2 && dev && sap_mop64(mud, L.sap_isolv, L.sap_nmr, L.sap_ncy, r, p, c))
   {
      // FP64 variant: the Schwarz cycles on FP64 fields
      if (nn_use)
      {
         diff_s2s(vt, 2, rho, chi);  // the residual of the cycles
         nn_correct(c, p);
         Dw((float)mud, phi, chi);
      }
      return;
   }
   set_s2zero(vt, 2, phi);
   assign_s2s(vt, 2, rho, chi);
   for (int n = 0; n < L.sap_ncy; n++) sap((float)mud, L.sap_isolv, L.sap_nmr, phi, chi);
   if (nn_use && dev)
   {
      // learned correction from the residual the Schwarz cycles leave
      nn_correct(c, p);
      Dw((float)mud, phi, chi);
      return;
   }
   diff_s2s(vt, 2, rho, chi);
}

extern "C" void oqh_set_gcr_orth(int mode)
{
   if (mode != 0 && mode != 1)
   {
      fail("oqh_set_gcr_orth", "the mode must be 0 (per-term) or 1 (two-pass modified Gram-Schmidt)");
      return;
   }
   lib().gcr_cgs = mode;
}

extern "C" double sap_gcr(int nkv, int nmx, int istop, double res, double mu, spinor_dble *eta,
                          spinor_dble *psi, int *ifail, int *status)
{
   Lib &L = lib();
   if (!require_geometry("sap_gcr [sap_gcr.c]")) return 0.0;
   if (L.sap_ncy == 0)
   {
      fail("sap_gcr [sap_gcr.c]", "SAP parameters are not set");
      return 0.0;
   }
   int vt = L.g.VT;
   ifail[0] = 0;
   status[0] = 0;
   double rho = 0.0;

   if (!sap_prepare(L.sap_isolv, "sap_gcr [sap_gcr.c]", true))
      ifail[0] = -3;
   else
   {
      double rho0 = unorm_dble(vt, 3, eta);
      if (rho0 != 0.0)
      {
         nn_use = nn_ready();
         const int nwsd = nn_use ? 4 : 2;
         alloc_ws(L.ws_used + 2 * nkv + 1);
         alloc_wsd(L.wsd_used + nwsd);
         spinor **ws = reserve_ws(2 * nkv + 1);
         spinor_dble **wsd = reserve_wsd(nwsd);
         spinor_dble *rsd = wsd[1];
         mud = mu;
         assign_sd2sd(vt, 2, eta, rsd);
         scale_dble(vt, 2, 1.0 / rho0, rsd);
         // with the learned preconditioner the solve may use half of the
         // Krylov vectors; if it has not converged then, plain SAP continues
         // from its solution with the rest, so that a call never uses more
         // than nmx vectors and status counts both parts
         const int nmx1 = nn_use ? (nmx > 1 ? nmx / 2 : 1) : nmx;
         rho = fgcr(vt, Dop, Mop, ws, wsd, nkv, nmx1, istop, res, rsd, psi, status);
         if (nn_use && status[0] < 0)
         {
            nn_use = false;
            status[0] = -1;
            if (nmx > nmx1)
            {
               // r2 = rsd - D psi, solve D dpsi = r2 to the original tolerance
               spinor_dble *r2 = wsd[2], *dpsi = wsd[3];
               Dw_dble(mu, psi, r2);
               combine_spinor_dble(vt, 2, r2, rsd, -1.0, 1.0);
               double n0, n2;
               if (istop)
               {
                  n0 = unorm_dble(vt, 3, rsd);
                  n2 = unorm_dble(vt, 3, r2);
               }
               else
               {
                  n0 = sqrt(norm_square_dble(vt, 3, rsd).q[0]);
                  n2 = sqrt(norm_square_dble(vt, 3, r2).q[0]);
               }
               if (!(n2 < n0))
               {
                  // the learned solve made no progress: plain SAP from psi = 0
                  set_sd2zero(vt, 2, psi);
                  assign_sd2sd(vt, 2, rsd, r2);
                  n2 = n0;
               }
               int st2 = 0;
               rho = fgcr(vt, Dop, Mop, ws, wsd, nkv, nmx - nmx1, istop, res * n0 / n2, r2, dpsi, &st2);
               mulr_spinor_add_dble(vt, 2, psi, dpsi, 1.0);
               status[0] = (st2 < 0) ? -1 : nmx1 + st2;
            }
         }
         nn_use = false;
         scale_dble(vt, 2, rho0, psi);
         rho *= rho0;
         if (status[0] < 0)
         {
            ifail[0] = -1;
            status[0] = nmx;
         }
         release_wsd();
         release_ws();
      }
      else
         set_sd2zero(vt, 2, psi);
   }

   if (ifail[0] <= -2)
   {
      if (psi != eta) set_sd2zero(vt, 2, psi);
      if (istop)
         rho = unorm_dble(vt, 3, eta);
      else
      {
         qflt q = norm_square_dble(vt, 3, eta);
         rho = sqrt(q.q[0]);
      }
   }
   return rho;
}

// ---------------------------------------------------------------------------
// several right-hand sides (no counterpart in the reference): the FGCR
// iterations of all sources advance together, the Schwarz cycles of all
// active sources run in the same launches (FP32 SAP, blocks of 256 points),
// each source keeps its own Krylov space, coefficients and restarts
// (fgcr.c:215-299 per source)
// ---------------------------------------------------------------------------

namespace {

struct MultiSrc
{
   spinor **phi = nullptr, *rho = nullptr, **chi = nullptr;
   spinor_dble *wrk = nullptr, *rsd = nullptr, *psi = nullptr;
   double rho0 = 0.0, rn = 0.0, tol = 0.0, rn_old = 0.0;
   int k = 0, status = 0;
   bool done = false, failed = false;
   std::vector<complex> a, c;
   std::vector<float> b;
   float2 *dcoef = nullptr, *gram = nullptr;
};

}  // namespace

extern "C" void sap_gcr_multi(int n, int nkv, int nmx, int istop, double res, double mu, spinor_dble **eta,
                              spinor_dble **psi, double *rho_out, int *ifail, int *status)
{
   const char *where = "sap_gcr_multi";
   Lib &L = lib();
   if (!require_geometry(where)) return;
   if (n < 1 || nkv < 1 || nmx < 1 || res <= DBL_EPSILON || istop < 0 || istop > 1)
   {
      fail(where, "Parameters are out of range");
      return;
   }
   if (L.sap_ncy == 0)
   {
      fail(where, "SAP parameters are not set");
      return;
   }
   const int vt = L.g.VT, V = L.g.V;
   // the joint path: FP32 SAP with the register kernel, 2-norm stopping
   // criterion, no network, two-pass Gram-Schmidt (oqh_set_gcr_orth(1));
   // otherwise one solve after the other
   bool joint = (L.sap_prec == 0 && istop == 0 && !nn_ready() && L.gcr_cgs == 1);
   if (joint && !sap_prepare(L.sap_isolv, where, true)) joint = false;
   int bs_vol = L.sap_bs[0] * L.sap_bs[1] * L.sap_bs[2] * L.sap_bs[3];
   if (bs_vol != 256) joint = false;
   if (!joint)
   {
      for (int s = 0; s < n; s++) rho_out[s] = sap_gcr(nkv, nmx, istop, res, mu, eta[s], psi[s], ifail + s, status + s);
      return;
   }
   const int SB_ = nkv, SC_ = nkv + 1, SR_ = nkv + 2, NC = nkv + 3;
   const double PRECISION_LIMIT = (double)(100.0f * FLT_EPSILON);
   alloc_ws(L.ws_used + n * (2 * nkv + 1));
   alloc_wsd(L.wsd_used + 2 * n);
   std::vector<MultiSrc> S(n);
   float2 *dcoef = (float2 *)dev_alloc(sizeof(float2) * NC * n);
   float2 *dgram = (float2 *)dev_alloc(sizeof(float2) * nkv * nkv * n);
   float2 *hcoef = nullptr;
   HIP_OK(hipHostMalloc((void **)&hcoef, sizeof(float2) * NC * n, 0));
   auto F = [](spinor *p) { return (float2 *)device_field(p, F32, &dummy_off); };
   mud = mu;
   for (int s = 0; s < n; s++)
   {
      MultiSrc &m = S[s];
      spinor **ws = reserve_ws(2 * nkv + 1);
      spinor_dble **wsd = reserve_wsd(2);
      m.phi = ws;
      m.rho = ws[nkv];
      m.chi = ws + nkv + 1;
      m.wrk = wsd[0];
      m.rsd = wsd[1];
      m.psi = psi[s];
      m.a.assign((size_t)nkv * nkv, complex{0.0f, 0.0f});
      m.c.assign(nkv, complex{0.0f, 0.0f});
      m.b.assign(nkv, 0.0f);
      m.dcoef = dcoef + (size_t)s * NC;
      m.gram = dgram + (size_t)s * nkv * nkv;
      ifail[s] = 0;
      status[s] = 0;
      rho_out[s] = 0.0;
      m.rho0 = unorm_dble(vt, 3, eta[s]);
      if (m.rho0 == 0.0)
      {
         set_sd2zero(vt, 2, psi[s]);
         m.done = true;
         continue;
      }
      assign_sd2sd(vt, 2, eta[s], m.rsd);
      scale_dble(vt, 2, 1.0 / m.rho0, m.rsd);
      set_sd2zero(vt, 2, m.psi);
      assign_sd2s(vt, 2, m.rsd, m.rho);
      m.rn = get_res(vt, 0, m.rho);
      m.tol = res * m.rn;
      m.rn_old = m.rn;
      if (m.rn <= m.tol) m.done = true;
   }
   std::vector<int> act;
   std::vector<float2 *> pl, rl, qv(nkv);
   while (true)
   {
      act.clear();
      for (int s = 0; s < n; s++)
         if (!S[s].done) act.push_back(s);
      if (act.empty()) break;
      // Mop of every active source: phi = SAP(rho), chi = rho - residual
      for (int s : act)
      {
         set_s2zero(vt, 2, S[s].phi[S[s].k]);
         assign_s2s(vt, 2, S[s].rho, S[s].chi[S[s].k]);
      }
      for (size_t i0 = 0; i0 < act.size(); i0 += 16)
      {
         pl.clear();
         rl.clear();
         for (size_t i = i0; i < act.size() && i < i0 + 16; i++)
         {
            pl.push_back(F(S[act[i]].phi[S[act[i]].k]));
            rl.push_back(F(S[act[i]].chi[S[act[i]].k]));
         }
         sap_multi_device((float)mud, L.sap_isolv, L.sap_nmr, L.sap_ncy, (int)pl.size(), pl.data(), rl.data());
      }
      for (int s : act) diff_s2s(vt, 2, S[s].rho, S[s].chi[S[s].k]);
      // Gram-Schmidt step of every active source, one read-back for all
      for (int s : act)
      {
         MultiSrc &m = S[s];
         for (int l = 0; l < m.k; l++) qv[l] = F(m.chi[l]);
         gs_classical(F(m.chi[m.k]), qv.data(), m.k, F(m.rho), m.dcoef, SB_, SC_, SR_, V, m.gram, nkv);
      }
      HIP_OK(hipMemcpyAsync(hcoef, dcoef, sizeof(float2) * NC * n, hipMemcpyDeviceToHost, L.stream));
      HIP_OK(hipStreamSynchronize(L.stream));
      for (int s : act)
      {
         MultiSrc &m = S[s];
         const float2 *hc = hcoef + (size_t)s * NC;
         const int k = m.k;
         for (int l = 0; l < k; l++)
         {
            m.a[(size_t)nkv * l + k].re = hc[l].x;
            m.a[(size_t)nkv * l + k].im = hc[l].y;
         }
         m.b[k] = hc[SB_].x;
         if (m.b[k] == 0.0f) fail("normalize [salg.c]", "Vector has vanishing norm");
         m.c[k].re = hc[SC_].x;
         m.c[k].im = hc[SC_].y;
         m.rn = sqrt((double)hc[SR_].y);
         status[s] += 1;
         if (!(m.rn <= m.tol || m.rn < PRECISION_LIMIT * m.rn_old || k == nkv - 1 || status[s] >= nmx))
         {
            m.k++;
            continue;
         }
         // restart: update_psi (fgcr.c:181-212)
         for (int l = k; l >= 0; l--)
         {
            float zr = m.c[l].re, zi = m.c[l].im;
            for (int i = l + 1; i <= k; i++)
            {
               zr -= (m.a[(size_t)l * nkv + i].re * m.c[i].re - m.a[(size_t)l * nkv + i].im * m.c[i].im);
               zi -= (m.a[(size_t)l * nkv + i].re * m.c[i].im + m.a[(size_t)l * nkv + i].im * m.c[i].re);
            }
            float r = 1.0f / m.b[l];
            m.c[l].re = zr * r;
            m.c[l].im = zi * r;
         }
         set_s2zero(vt, 2, m.rho);
         for (int l = k; l >= 0; l--) mulc_spinor_add(vt, 2, m.rho, m.phi[l], m.c[l]);
         add_s2sd(vt, 2, m.rho, m.psi);
         Dw_dble(mud, m.psi, m.wrk);
         diff_sd2s(vt, 2, m.rsd, m.wrk, m.rho);
         m.rn = get_res(vt, 0, m.rho);
         m.k = 0;
         m.rn_old = m.rn;
         if (status[s] >= nmx && m.rn > m.tol)
         {
            m.failed = true;
            m.done = true;
         }
         else if (m.rn <= m.tol)
            m.done = true;
      }
   }
   for (int s = 0; s < n; s++)
   {
      MultiSrc &m = S[s];
      if (m.rho0 == 0.0) continue;
      scale_dble(vt, 2, m.rho0, m.psi);
      rho_out[s] = m.rn * m.rho0;
      if (m.failed)
      {
         ifail[s] = -1;
         status[s] = nmx;
      }
   }
   HIP_OK(hipHostFree(hcoef));
   dev_free(dcoef);
   dev_free(dgram);
   for (int s = 0; s < n; s++)
   {
      release_wsd();
      release_ws();
   }
}

