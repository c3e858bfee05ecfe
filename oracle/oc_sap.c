/*
 * oc_sap.c -- oracle: SAP block grid, block operators, block MinRes, the
 * Schwarz sweep, flexible GCR and sap_gcr.  TEST INFRASTRUCTURE ONLY.
 *
 *   block grid order ........ block/blk_grid.c:63-113 (colour 0 blocks
 *                             first, lexicographic block coordinates)
 *   block-local site order .. lattice/blk_geometry.c:56-170 (cache blocks,
 *                             even first; neighbour outside block = vol)
 *   Dw_blk, Dwhat_blk, Dwoo_blk, Dwoe_blk, Dweo_blk
 *                             dirac/Dw.c:1475-1907 (Dirichlet exterior;
 *                             open-bc boundary points ibp zeroed)
 *   blk_mres / blk_eo_mres .. sap/blk_solv.c:1117-1186
 *   sap() ................... sap/sap.c:592-790 (update_flds0/1/2)
 *   Dw_bnd + sap_com ........ dirac/Dw_bnd.c:801-1011, sap/sap_com.c:465-501:
 *                             rho(y) -= hopping term from the block correction
 *                             for every exterior neighbour y of the block
 *   fgcr .................... linsolv/fgcr.c:128-299
 *   sap_gcr ................. sap/sap_gcr.c:120-254
 *   block sw field .......... block/map_sw2blk.c:53-100 (odd part inverted in
 *                             FP64 when isolv=1, then rounded to FP32)
 * Supported boundary conditions: 0 (open) and 3 (periodic).
 */
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <float.h>
#include "oc_internal.h"

void oc_hop_term_f(int mu, int sgn, int dag, const float *u, const float *v, float coef, float *r);
void oc_hop_term_d(int mu, int sgn, int dag, const double *u, const double *v, double coef, double *r);

static const int *blk_cbs(const int *bs, int *cbs)
{
   cbs[0] = bs[0];
   for (int mu = 1; mu < 4; mu++)
      cbs[mu] = (bs[mu] % 4 == 0) ? 4 : ((bs[mu] % 3 == 0) ? 3 : 2);
   return cbs;
}

void oc_build_blocks(oc_ctx *c)
{
   int *bs = c->sap_bs, nbl[4], cbs[4], nbs[4];
   if (c->blk)
   {
      for (int n = 0; n < c->nb; n++)
      {
         free(c->blk[n].ipt); free(c->blk[n].iup); free(c->blk[n].idn); free(c->blk[n].imb);
      }
      free(c->blk);
   }
   for (int mu = 0; mu < 4; mu++) nbl[mu] = c->L[mu] / bs[mu];
   c->nb = nbl[0] * nbl[1] * nbl[2] * nbl[3];
   c->nbh = c->nb / 2;
   c->bvol = bs[0] * bs[1] * bs[2] * bs[3];
   c->blk = calloc(c->nb, sizeof(oc_block));
   blk_cbs(bs, cbs);
   for (int mu = 0; mu < 4; mu++) nbs[mu] = bs[mu] / cbs[mu];

   int ib = 0, m[4];
   for (int col = 0; col < 2; col++)
      for (m[0] = 0; m[0] < nbl[0]; m[0]++)
      for (m[1] = 0; m[1] < nbl[1]; m[1]++)
      for (m[2] = 0; m[2] < nbl[2]; m[2]++)
      for (m[3] = 0; m[3] < nbl[3]; m[3]++)
      {
         if (((m[0] + m[1] + m[2] + m[3]) & 1) != col) continue;
         oc_block *b = c->blk + ib++;
         int vol = c->bvol, ofs[2] = {0, vol / 2}, j[4], x[4], y[4];
         b->vol = vol;
         for (int mu = 0; mu < 4; mu++) b->bo[mu] = m[mu] * bs[mu];
         b->ipt = malloc(sizeof(int) * vol);
         b->iup = malloc(sizeof(int[4]) * vol);
         b->idn = malloc(sizeof(int[4]) * vol);
         b->imb = malloc(sizeof(int) * (vol + 1));
         for (j[0] = 0; j[0] < nbs[0]; j[0]++)
         for (j[1] = 0; j[1] < nbs[1]; j[1]++)
         for (j[2] = 0; j[2] < nbs[2]; j[2]++)
         for (j[3] = 0; j[3] < nbs[3]; j[3]++)
            for (x[0] = 0; x[0] < cbs[0]; x[0]++)
            for (x[1] = 0; x[1] < cbs[1]; x[1]++)
            for (x[2] = 0; x[2] < cbs[2]; x[2]++)
            for (x[3] = 0; x[3] < cbs[3]; x[3]++)
            {
               for (int mu = 0; mu < 4; mu++) y[mu] = j[mu] * cbs[mu] + x[mu];
               int bl = y[3] + bs[3] * (y[2] + bs[2] * (y[1] + bs[1] * y[0]));
               b->ipt[bl] = ofs[(y[0] + y[1] + y[2] + y[3]) & 1]++;
            }
         for (y[0] = 0; y[0] < bs[0]; y[0]++)
         for (y[1] = 0; y[1] < bs[1]; y[1]++)
         for (y[2] = 0; y[2] < bs[2]; y[2]++)
         for (y[3] = 0; y[3] < bs[3]; y[3]++)
         {
            int bl = y[3] + bs[3] * (y[2] + bs[2] * (y[1] + bs[1] * y[0]));
            int ix = b->ipt[bl], g[4];
            for (int mu = 0; mu < 4; mu++) g[mu] = b->bo[mu] + y[mu];
            b->imb[ix] = c->ipt[g[3] + c->L[3] * (g[2] + c->L[2] * (g[1] + c->L[1] * g[0]))];
            for (int mu = 0; mu < 4; mu++)
            {
               int z[4];
               memcpy(z, y, sizeof(z));
               z[mu] = y[mu] + 1;
               b->iup[ix][mu] = (z[mu] < bs[mu])
                  ? b->ipt[z[3] + bs[3] * (z[2] + bs[2] * (z[1] + bs[1] * z[0]))] : vol;
               z[mu] = y[mu] - 1;
               b->idn[ix][mu] = (z[mu] >= 0)
                  ? b->ipt[z[3] + bs[3] * (z[2] + bs[2] * (z[1] + bs[1] * z[0]))] : vol;
            }
         }
         b->imb[vol] = b->imb[0];
      }
}

int oc_sap_nblocks(oc_ctx *c)
{
   if (!c->blk) oc_build_blocks(c);
   return c->nb;
}

void oc_sap_block_info(oc_ctx *c, int n, int *bo, int *vol)
{
   if (!c->blk) oc_build_blocks(c);
   memcpy(bo, c->blk[n].bo, sizeof(int) * 4);
   *vol = c->blk[n].vol;
}

void oc_sap_block_imb(oc_ctx *c, int n, int *imb)
{
   if (!c->blk) oc_build_blocks(c);
   memcpy(imb, c->blk[n].imb, sizeof(int) * c->blk[n].vol);
}

int oc_sap_prepare(oc_ctx *c)
{
   int ifail = 0;
   oc_build_blocks(c);
   oc_assign_ud2u(c);
   if (c->swd_inv[0] || c->swd_inv[1]) oc_sw_term(c, 3);
   free(c->sw_bgr);
   free(c->swd_bgr);
   c->sw_bgr = malloc(sizeof(float) * 72 * (size_t)c->V);
   c->swd_bgr = malloc(sizeof(double) * 72 * (size_t)c->V);
   for (int ix = 0; ix < c->V; ix++)
   {
      const double *p = c->swd + (size_t)72 * ix;
      double ms[72];
      if (c->sap_isolv && ix >= c->V / 2)
      {
         ifail |= oc_inv_pauli_dble(0.0, p, ms);
         ifail |= oc_inv_pauli_dble(0.0, p + 36, ms + 36);
         p = ms;
      }
      for (int i = 0; i < 72; i++)
      {
         c->sw_bgr[(size_t)72 * ix + i] = (float)p[i];
         c->swd_bgr[(size_t)72 * ix + i] = p[i];
      }
   }
   c->sap_ready = 1;
   return ifail;
}

/* ---- block-local helpers (fields have vol+1 spinors, last one zero) ---- */

static int blk_bnd_pt(const oc_ctx *c, const oc_block *b, int x)
{
   int t = c->x0[b->imb[x]];
   if (c->bc == 3) return 0;
   return (t == 0) || (t == c->L[0] - 1 && c->bc == 0);
}

static unsigned char *bnd_mask(const oc_ctx *c, const oc_block *b)
{
   unsigned char *m = calloc(b->vol + 1, 1);
   for (int x = 0; x < b->vol; x++) m[x] = (unsigned char)blk_bnd_pt(c, b, x);
   return m;
}

/* ---- FP64 BLAS-1 of the FP64 instance: plain double sums ---- */

static double nsq_d(int n, const double *s)
{
   double x = 0.0;
   for (size_t i = 0; i < (size_t)24 * n; i++) x += s[i] * s[i];
   return x;
}

static void sprod_d(int n, const double *s, const double *r, double *z)
{
   double vr = 0.0, vi = 0.0;
   for (size_t i = 0; i < (size_t)12 * n; i++)
   {
      vr += s[2 * i] * r[2 * i] + s[2 * i + 1] * r[2 * i + 1];
      vi += s[2 * i] * r[2 * i + 1] - s[2 * i + 1] * r[2 * i];
   }
   z[0] = vr;
   z[1] = vi;
}

#define R float
#define SFX
#define UFLD(c) ((c)->u)
#define SWB(c) ((c)->sw_bgr)
#define MULP oc_mul_pauli
#define HOP oc_hop_term_f
#define NSQ oc_norm_square
#define SPROD oc_spinor_prod
#define MULCADD oc_mulc_spinor_add
#define RMIN FLT_MIN
#include "oc_sap_tmpl.h"
#undef R
#undef SFX
#undef UFLD
#undef SWB
#undef MULP
#undef HOP
#undef NSQ
#undef SPROD
#undef MULCADD
#undef RMIN

#define R double
#define SFX _dble
#define UFLD(c) ((c)->ud)
#define SWB(c) ((c)->swd_bgr)
#define MULP oc_mul_pauli_dble
#define HOP oc_hop_term_d
#define NSQ nsq_d
#define SPROD sprod_d
#define MULCADD oc_mulc_spinor_add_dble
#define RMIN DBL_MIN
#include "oc_sap_tmpl.h"
#undef R
#undef SFX
#undef UFLD
#undef SWB
#undef MULP
#undef HOP
#undef NSQ
#undef SPROD
#undef MULCADD
#undef RMIN

#define ZS(p, x) memset((p) + 24 * (size_t)(x), 0, 24 * sizeof(float))

/* ------------------------------------------------------------------ */
/* flexible GCR and sap_gcr                                             */
/* ------------------------------------------------------------------ */

static double get_res(int V, int istop, const float *rho)
{
   if (istop) return (double)oc_unorm(V, rho);
   return sqrt((double)oc_norm_square(V, rho));
}

double oc_sap_gcr(oc_ctx *c, int nkv, int nmx, int istop, double res, double mu,
                  double *eta, double *psi, int *ifail, int *status)
{
   int V = c->V;
   size_t NS = (size_t)24 * V;
   ifail[0] = 0;
   status[0] = 0;
   c->sap_ready = 0;
   if (oc_sap_prepare(c))
   {
      ifail[0] = -3;
      if (psi != eta) memset(psi, 0, NS * sizeof(double));
      if (istop) return oc_unorm_dble(V, eta);
      double q[2];
      oc_norm_square_dble(V, eta, q);
      return sqrt(q[0]);
   }
   double rho0 = oc_unorm_dble(V, eta);
   if (rho0 == 0.0)
   {
      memset(psi, 0, NS * sizeof(double));
      return 0.0;
   }
   double *rsd = malloc(NS * sizeof(double)), *wrk = malloc(NS * sizeof(double));
   for (size_t i = 0; i < NS; i++) rsd[i] = eta[i] * (1.0 / rho0);
   float **phi = malloc(sizeof(float *) * nkv), **chi = malloc(sizeof(float *) * nkv);
   for (int k = 0; k < nkv; k++)
   {
      phi[k] = calloc(NS, sizeof(float));
      chi[k] = calloc(NS, sizeof(float));
   }
   float *rho = malloc(NS * sizeof(float));
   float (*a)[2] = calloc((size_t)nkv * nkv, sizeof(float[2]));
   float *b = calloc(nkv, sizeof(float));
   float (*cc)[2] = calloc(nkv, sizeof(float[2]));

   memset(psi, 0, NS * sizeof(double));
   for (size_t i = 0; i < NS; i++) rho[i] = (float)rsd[i];
   double rn = get_res(V, istop, rho), tol = res * rn, rn_old;
   float muf = (float)mu;

   while (rn > tol)
   {
      int k;
      rn_old = rn;
      for (k = 0;; k++)
      {
         /* Mop: phi = M rho, chi = D phi (through the SAP residual) */
         memset(phi[k], 0, NS * sizeof(float));
         memcpy(chi[k], rho, NS * sizeof(float));
         for (int n = 0; n < c->sap_ncy; n++)
            oc_sap(c, muf, c->sap_isolv, c->sap_nmr, phi[k], chi[k]);
         for (size_t i = 0; i < NS; i++) chi[k][i] = rho[i] - chi[k][i];
         for (int l = 0; l < k; l++)
         {
            oc_spinor_prod(V, chi[l], chi[k], a[nkv * l + k]);
            oc_mulc_spinor_add(V, chi[k], chi[l], -a[nkv * l + k][0], -a[nkv * l + k][1]);
         }
         float r = (float)sqrt((double)oc_norm_square(V, chi[k]));
         b[k] = r;
         if (r != 0.0f)
            for (size_t i = 0; i < NS; i++) chi[k][i] *= 1.0f / r;
         oc_spinor_prod(V, chi[k], rho, cc[k]);
         oc_mulc_spinor_add(V, rho, chi[k], -cc[k][0], -cc[k][1]);
         rn = get_res(V, istop, rho);
         status[0] += 1;
         if (rn <= tol || rn < (double)(100.0f * FLT_EPSILON) * rn_old || k == nkv - 1 ||
             status[0] >= nmx)
            break;
      }
      /* update_psi */
      for (int l = k; l >= 0; l--)
      {
         float zr = cc[l][0], zi = cc[l][1];
         for (int i = l + 1; i <= k; i++)
         {
            zr -= (a[l * nkv + i][0] * cc[i][0] - a[l * nkv + i][1] * cc[i][1]);
            zi -= (a[l * nkv + i][0] * cc[i][1] + a[l * nkv + i][1] * cc[i][0]);
         }
         float r = 1.0f / b[l];
         cc[l][0] = zr * r;
         cc[l][1] = zi * r;
      }
      memset(rho, 0, NS * sizeof(float));
      for (int l = k; l >= 0; l--) oc_mulc_spinor_add(V, rho, phi[l], cc[l][0], cc[l][1]);
      for (size_t i = 0; i < NS; i++) psi[i] += (double)rho[i];
      oc_Dw_dble(c, mu, psi, wrk);
      for (size_t i = 0; i < NS; i++) rho[i] = (float)(rsd[i] - wrk[i]);
      rn = get_res(V, istop, rho);
      if (status[0] >= nmx && rn > tol)
      {
         status[0] = -1;
         break;
      }
   }
   for (size_t i = 0; i < NS; i++) psi[i] *= rho0;
   rn *= rho0;
   if (status[0] < 0)
   {
      ifail[0] = -1;
      status[0] = nmx;
   }
   for (int k = 0; k < nkv; k++)
   {
      free(phi[k]);
      free(chi[k]);
   }
   free(phi); free(chi); free(rho); free(a); free(b); free(cc); free(rsd); free(wrk);
   return rn;
}
