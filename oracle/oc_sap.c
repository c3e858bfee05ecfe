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
   c->sw_bgr = malloc(sizeof(float) * 72 * (size_t)c->V);
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
      for (int i = 0; i < 72; i++) c->sw_bgr[(size_t)72 * ix + i] = (float)p[i];
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

static void blk_pauli(const oc_ctx *c, const oc_block *b, int x, float mu, const float *s, float *r)
{
   const float *m = c->sw_bgr + (size_t)72 * b->imb[x];
   oc_mul_pauli(mu, m, s, r);
   oc_mul_pauli(-mu, m + 36, s + 12, r + 12);
}

/* r[x] += coef * hopping terms from in-block neighbours of x */
static void blk_hop(const oc_ctx *c, const oc_block *b, int x, const float *s,
                    const unsigned char *skip, float coef, float *r)
{
   int gx = b->imb[x];
   for (int mu = 0; mu < 4; mu++)
   {
      int y = b->iup[x][mu];
      if (y < b->vol && !(skip && skip[y]))
         oc_hop_term_f(mu, +1, 0, c->u + (size_t)18 * oc_link_fwd(c, gx, mu), s + 24 * (size_t)y, coef, r);
      y = b->idn[x][mu];
      if (y < b->vol && !(skip && skip[y]))
         oc_hop_term_f(mu, -1, 1, c->u + (size_t)18 * oc_link_bwd(c, gx, mu), s + 24 * (size_t)y, coef, r);
   }
}

static unsigned char *bnd_mask(const oc_ctx *c, const oc_block *b)
{
   unsigned char *m = calloc(b->vol + 1, 1);
   for (int x = 0; x < b->vol; x++) m[x] = (unsigned char)blk_bnd_pt(c, b, x);
   return m;
}

#define ZS(p, x) memset((p) + 24 * (size_t)(x), 0, 24 * sizeof(float))

void oc_Dw_blk(oc_ctx *c, int n, float mu, const float *sin, float *r)
{
   oc_block *b = c->blk + n;
   int vol = b->vol, h = vol / 2;
   float muo = c->eoflg ? 0.0f : mu;
   float *s = malloc(sizeof(float) * 24 * (vol + 1));
   memcpy(s, sin, sizeof(float) * 24 * vol);
   ZS(s, vol);
   unsigned char *bp = bnd_mask(c, b);
   for (int x = 0; x < vol; x++)
      if (bp[x]) ZS(s, x);
   for (int x = 0; x < vol; x++)
   {
      float *rr = r + 24 * (size_t)x;
      if (bp[x]) { ZS(r, x); continue; }
      blk_pauli(c, b, x, (x < h) ? mu : muo, s + 24 * (size_t)x, rr);
      blk_hop(c, b, x, s, NULL, -0.5f, rr);
   }
   ZS(r, vol);
   free(bp);
   free(s);
}

/* even-site Schur complement on the block; sw_bgr odd part must be inverted */
void oc_Dwhat_blk(oc_ctx *c, int n, float mu, const float *sin, float *r)
{
   oc_block *b = c->blk + n;
   int vol = b->vol, h = vol / 2;
   float *s = calloc((size_t)24 * (vol + 1), sizeof(float));
   float *t = calloc((size_t)24 * (vol + 1), sizeof(float));
   memcpy(s, sin, sizeof(float) * 24 * vol);
   unsigned char *bp = bnd_mask(c, b);
   for (int x = 0; x < h; x++) blk_pauli(c, b, x, mu, s + 24 * (size_t)x, r + 24 * (size_t)x);
   for (int x = 0; x < h; x++)
      if (bp[x]) ZS(s, x);
   for (int x = h; x < vol; x++)
   {
      if (bp[x]) continue;
      float acc[24] = {0};
      blk_hop(c, b, x, s, NULL, -0.5f, acc);
      blk_pauli(c, b, x, 0.0f, acc, t + 24 * (size_t)x);
   }
   for (int x = 0; x < h; x++) blk_hop(c, b, x, t, bp, 0.5f, r + 24 * (size_t)x);
   for (int x = 0; x < h; x++)
      if (bp[x]) ZS(r, x);
   free(bp);
   free(s);
   free(t);
}

static void blk_Dwoo(oc_ctx *c, const oc_block *b, float mu, float *s, float *r)
{
   if (c->eoflg) mu = 0.0f;
   for (int x = b->vol / 2; x < b->vol; x++)
   {
      if (blk_bnd_pt(c, b, x)) { ZS(s, x); ZS(r, x); continue; }
      float out[24];
      blk_pauli(c, b, x, mu, s + 24 * (size_t)x, out);
      memcpy(r + 24 * (size_t)x, out, sizeof(out));
   }
}

static void blk_Dwoe(oc_ctx *c, const oc_block *b, float *s, float *r)
{
   int h = b->vol / 2;
   for (int x = 0; x < h; x++)
      if (blk_bnd_pt(c, b, x)) ZS(s, x);
   for (int x = h; x < b->vol; x++)
   {
      float acc[24] = {0};
      if (!blk_bnd_pt(c, b, x)) blk_hop(c, b, x, s, NULL, -0.5f, acc);
      memcpy(r + 24 * (size_t)x, acc, sizeof(acc));
   }
}

static void blk_Dweo(oc_ctx *c, const oc_block *b, float *s, float *r)
{
   int h = b->vol / 2;
   unsigned char *bp = bnd_mask(c, b);
   for (int x = h; x < b->vol; x++)
      if (bp[x]) ZS(s, x);
   for (int x = 0; x < h; x++) blk_hop(c, b, x, s, bp, 0.5f, r + 24 * (size_t)x);
   for (int x = 0; x < h; x++)
      if (bp[x]) ZS(r, x);
   free(bp);
}

static void blk_Dwee(oc_ctx *c, const oc_block *b, float mu, float *s, float *r)
{
   for (int x = 0; x < b->vol / 2; x++)
   {
      if (blk_bnd_pt(c, b, x)) { ZS(s, x); ZS(r, x); continue; }
      float out[24];
      blk_pauli(c, b, x, mu, s + 24 * (size_t)x, out);
      memcpy(r + 24 * (size_t)x, out, sizeof(out));
   }
}

/* Dwee_blk, Dwoo_blk, Dwoe_blk, Dweo_blk (dirac/Dw.c:1584-1826) on copies
   of the block fields b.s[k] = s and b.s[l] = r (vol points each, block
   order); s and r are written back, with the zeros the operators put on the
   boundary points of open bc.  op: 0 Dwee, 1 Dwoo, 2 Dwoe, 3 Dweo. */
void oc_blk_eo_op(oc_ctx *c, int n, int op, float mu, float *s, float *r)
{
   oc_block *b = c->blk + n;
   size_t nb = sizeof(float) * 24 * (size_t)b->vol;
   float *sc = calloc((size_t)24 * (b->vol + 1), sizeof(float));
   float *rc = calloc((size_t)24 * (b->vol + 1), sizeof(float));
   memcpy(sc, s, nb);
   memcpy(rc, r, nb);
   if (op == 0) blk_Dwee(c, b, mu, sc, rc);
   else if (op == 1) blk_Dwoo(c, b, mu, sc, rc);
   else if (op == 2) blk_Dwoe(c, b, sc, rc);
   else blk_Dweo(c, b, sc, rc);
   memcpy(s, sc, nb);
   memcpy(r, rc, nb);
   free(sc);
   free(rc);
}

static void mres(oc_ctx *c, int n, float mu, int nmr, int eo, float *s0, float *s1)
{
   int vol = c->blk[n].vol, nv = eo ? vol / 2 : vol;
   float *s2 = calloc((size_t)24 * (vol + 1), sizeof(float));
   memset(s0, 0, sizeof(float) * 24 * nv);
   for (int imr = 0; imr < nmr; imr++)
   {
      if (eo) oc_Dwhat_blk(c, n, mu, s1, s2);
      else oc_Dw_blk(c, n, mu, s1, s2);
      float r = oc_norm_square(nv, s2), z[2];
      if (r < 2.0f * FLT_MIN) break;
      oc_spinor_prod(nv, s2, s1, z);
      r = 1.0f / r;
      z[0] *= r;
      z[1] *= r;
      oc_mulc_spinor_add(nv, s0, s1, z[0], z[1]);
      oc_mulc_spinor_add(nv, s1, s2, -z[0], -z[1]);
   }
   free(s2);
}

void oc_blk_mres(oc_ctx *c, int n, float mu, int nmr, float *s0, float *s1)
{
   mres(c, n, mu, nmr, 0, s0, s1);
}

/* rho(y) -= hopping term from s0 for every exterior neighbour y */
static void blk_bnd_update(oc_ctx *c, const oc_block *b, const float *s0, float *rho)
{
   for (int x = 0; x < b->vol; x++)
   {
      int gx = b->imb[x];
      for (int mu = 0; mu < 4; mu++)
      {
         if (b->idn[x][mu] == b->vol)
         {
            /* y = x - mu receives -1/2 (1-gamma_mu) U(y,mu) s0(x) */
            int gy = c->idn[gx][mu];
            float t[24] = {0};
            oc_hop_term_f(mu, +1, 0, c->u + (size_t)18 * oc_link_bwd(c, gx, mu), s0 + 24 * (size_t)x, -0.5f, t);
            for (int i = 0; i < 24; i++) rho[24 * (size_t)gy + i] -= t[i];
         }
         if (b->iup[x][mu] == b->vol)
         {
            /* y = x + mu receives -1/2 (1+gamma_mu) U(x,mu)^dag s0(x) */
            int gy = c->iup[gx][mu];
            float t[24] = {0};
            oc_hop_term_f(mu, -1, 1, c->u + (size_t)18 * oc_link_fwd(c, gx, mu), s0 + 24 * (size_t)x, -0.5f, t);
            for (int i = 0; i < 24; i++) rho[24 * (size_t)gy + i] -= t[i];
         }
      }
   }
}

void oc_sap(oc_ctx *c, float mu, int isolv, int nmr, float *psi, float *rho)
{
   if (!c->sap_ready) oc_sap_prepare(c);
   int vol = c->bvol, h = vol / 2;
   float *s0 = calloc((size_t)24 * (vol + 1) * c->nb, sizeof(float));
   float *s1 = calloc((size_t)24 * (vol + 1), sizeof(float));
   for (int ic = 0; ic < 2; ic++)
   {
      int n0 = ic ? c->nbh : 0, n1 = ic ? c->nb : c->nbh;
      for (int n = n0; n < n1; n++)
      {
         oc_block *b = c->blk + n;
         float *b0 = s0 + (size_t)24 * (vol + 1) * n;
         for (int x = 0; x < vol; x++)
            memcpy(s1 + 24 * (size_t)x, rho + 24 * (size_t)b->imb[x], 24 * sizeof(float));
         ZS(s1, vol);
         if (isolv == 0)
         {
            mres(c, n, mu, nmr, 0, b0, s1);
            for (int x = 0; x < vol; x++)
            {
               float *p = psi + 24 * (size_t)b->imb[x], *q = rho + 24 * (size_t)b->imb[x];
               for (int i = 0; i < 24; i++) p[i] += b0[24 * (size_t)x + i];
               memcpy(q, s1 + 24 * (size_t)x, 24 * sizeof(float));
            }
         }
         else
         {
            blk_Dwoo(c, b, 0.0f, s1, s1);
            blk_Dweo(c, b, s1, s1);
            mres(c, n, mu, nmr, 1, b0, s1);
            blk_Dwoe(c, b, b0, b0);
            blk_Dwoo(c, b, 0.0f, b0, b0);
            for (int x = 0; x < vol; x++)
            {
               float *p = psi + 24 * (size_t)b->imb[x], *q = rho + 24 * (size_t)b->imb[x];
               float *a = b0 + 24 * (size_t)x, *r1 = s1 + 24 * (size_t)x;
               if (x < h)
               {
                  for (int i = 0; i < 24; i++) p[i] += a[i];
                  memcpy(q, r1, 24 * sizeof(float));
               }
               else
               {
                  for (int i = 0; i < 24; i++)
                  {
                     a[i] = r1[i] - a[i];
                     p[i] += a[i];
                  }
                  if (c->eoflg)
                     memset(q, 0, 24 * sizeof(float));
                  else
                     for (int i = 0; i < 12; i++)
                     {
                        /* rho = -i mu gamma5 s0: spins 1,2 get -i mu, spins 3,4 +i mu */
                        float cm = (i < 6) ? -mu : mu;
                        q[2 * i] = -cm * a[2 * i + 1];
                        q[2 * i + 1] = cm * a[2 * i];
                     }
               }
            }
         }
      }
      for (int n = n0; n < n1; n++)
         blk_bnd_update(c, c->blk + n, s0 + (size_t)24 * (vol + 1) * n, rho);
   }
   free(s0);
   free(s1);
}

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
