/*
 * oc_core.c -- oracle: context, geometry, RANLUX, random fields, boundary
 * conditions.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Reference behaviour restated here:
 *   site order ipt ............ lattice/geometry.c:50-190 (thread blocks,
 *                               2x2x2x2 sub-blocks, cache blocks, even/odd)
 *   neighbours iup/idn ........ lattice/geogen.c (periodic at NPROC=1)
 *   RANLUX engine ............. random/ranlux_common.c:73-133, 470-520
 *   ranlxs / ranlxd streams ... random/ranlxs.c, random/ranlxd.c
 *   per-thread seeding ........ random/ranlux.c:136-171
 *   gauss / gauss_dble ........ random/gauss.c
 *   random_su3_dble ........... su3fcts/random_su3.c:125-181
 *   random_ud ................. uflds/uflds.c:159-222
 *   open bc / set_ud_phase .... lattice/bcnds.c:158-212,489-515;
 *                               uflds/uflds.c:325-376 (theta = 0 only)
 *   random_sd / random_s ...... sflds/sflds.c:1119-1267
 *   bnd_sd2zero ............... lattice/bcnds.c:275-354, 750-788
 *   unorm / unorm_dble ........ sflds/unorm.c:261-314
 */
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <stdio.h>
#include "oc_internal.h"

/* ------------------------------------------------------------------ */
/* geometry                                                             */
/* ------------------------------------------------------------------ */

static void cbs_of(const int *bs, int *cbs)
{
   cbs[0] = bs[0];
   for (int mu = 1; mu < 4; mu++)
      cbs[mu] = (bs[mu] % 4 == 0) ? 4 : ((bs[mu] % 3 == 0) ? 3 : 2);
}

static int lex_of(const int *L, const int *x)
{
   return x[3] + L[3] * (x[2] + L[2] * (x[1] + L[1] * x[0]));
}

/* number the sites of box [org, org+ext) in cache-block order */
static void number_box(oc_ctx *c, const int *org, const int *ext, int *ofs)
{
   int cbs[4], nbs[4], y[4], j[4], x[4];
   cbs_of(ext, cbs);
   for (int mu = 0; mu < 4; mu++) nbs[mu] = ext[mu] / cbs[mu];
   for (j[0] = 0; j[0] < nbs[0]; j[0]++)
   for (j[1] = 0; j[1] < nbs[1]; j[1]++)
   for (j[2] = 0; j[2] < nbs[2]; j[2]++)
   for (j[3] = 0; j[3] < nbs[3]; j[3]++)
      for (x[0] = 0; x[0] < cbs[0]; x[0]++)
      for (x[1] = 0; x[1] < cbs[1]; x[1]++)
      for (x[2] = 0; x[2] < cbs[2]; x[2]++)
      for (x[3] = 0; x[3] < cbs[3]; x[3]++)
      {
         for (int mu = 0; mu < 4; mu++) y[mu] = org[mu] + j[mu] * cbs[mu] + x[mu];
         int p = (y[0] + y[1] + y[2] + y[3]) & 1;
         c->ipt[lex_of(c->L, y)] = ofs[p]++;
      }
}

oc_ctx *oc_create(const int *L, const int *Ltrd)
{
   oc_ctx *c = calloc(1, sizeof(*c));
   int sbs[2][4], i[4], org[4], ext[4], ofs[2], nt[4];

   for (int mu = 0; mu < 4; mu++)
   {
      c->L[mu] = L[mu];
      c->Ltrd[mu] = Ltrd[mu];
      if (Ltrd[mu] < 4 || Ltrd[mu] % 2 || L[mu] % Ltrd[mu])
      {
         free(c);
         return NULL;
      }
      sbs[0][mu] = Ltrd[mu] / 2 + (Ltrd[mu] / 2) % 2;
      sbs[1][mu] = Ltrd[mu] - sbs[0][mu];
      nt[mu] = L[mu] / Ltrd[mu];
   }
   c->V = L[0] * L[1] * L[2] * L[3];
   c->VT = Ltrd[0] * Ltrd[1] * Ltrd[2] * Ltrd[3];
   c->nthread = c->V / c->VT;
   c->ipt = malloc(sizeof(int) * c->V);
   c->lex = malloc(sizeof(int) * c->V);
   c->iup = malloc(sizeof(int[4]) * c->V);
   c->idn = malloc(sizeof(int[4]) * c->V);
   c->x0 = malloc(sizeof(int) * c->V);

   /* thread block 0: 16 sub-blocks, each split into cache blocks */
   ofs[0] = 0;
   ofs[1] = c->V / 2;
   for (i[0] = 0; i[0] < 2; i[0]++)
   for (i[1] = 0; i[1] < 2; i[1]++)
   for (i[2] = 0; i[2] < 2; i[2]++)
   for (i[3] = 0; i[3] < 2; i[3]++)
   {
      for (int mu = 0; mu < 4; mu++)
      {
         ext[mu] = sbs[i[mu]][mu];
         org[mu] = i[mu] * sbs[0][mu];
      }
      number_box(c, org, ext, ofs);
   }
   /* other thread blocks: translated copies, offset by k*VT/2 per parity */
   for (int k = 1; k < c->nthread; k++)
   {
      int n = k, b[4], y[4], x[4];
      b[3] = n % nt[3]; n /= nt[3];
      b[2] = n % nt[2]; n /= nt[2];
      b[1] = n % nt[1]; n /= nt[1];
      b[0] = n;
      for (x[0] = 0; x[0] < Ltrd[0]; x[0]++)
      for (x[1] = 0; x[1] < Ltrd[1]; x[1]++)
      for (x[2] = 0; x[2] < Ltrd[2]; x[2]++)
      for (x[3] = 0; x[3] < Ltrd[3]; x[3]++)
      {
         for (int mu = 0; mu < 4; mu++) y[mu] = x[mu] + b[mu] * Ltrd[mu];
         c->ipt[lex_of(L, y)] = c->ipt[lex_of(L, x)] + k * (c->VT / 2);
      }
   }
   for (int lx = 0; lx < c->V; lx++)
      c->lex[c->ipt[lx]] = lx;

   for (int ix = 0; ix < c->V; ix++)
   {
      int lx = c->lex[ix], x[4], y[4];
      x[3] = lx % L[3];
      x[2] = (lx / L[3]) % L[2];
      x[1] = (lx / (L[3] * L[2])) % L[1];
      x[0] = lx / (L[3] * L[2] * L[1]);
      c->x0[ix] = x[0];
      for (int mu = 0; mu < 4; mu++)
      {
         memcpy(y, x, sizeof(y));
         y[mu] = (x[mu] + 1) % L[mu];
         c->iup[ix][mu] = c->ipt[lex_of(L, y)];
         y[mu] = (x[mu] + L[mu] - 1) % L[mu];
         c->idn[ix][mu] = c->ipt[lex_of(L, y)];
      }
   }

   c->bc = 3;
   c->cF[0] = c->cF[1] = 1.0;
   c->m0 = 0.0;
   c->csw = 1.0;
   c->ud = calloc((size_t)4 * c->V * 18, sizeof(double));
   c->u = calloc((size_t)4 * c->V * 18, sizeof(float));
   c->swd = calloc((size_t)2 * c->V * 36, sizeof(double));
   c->sw = calloc((size_t)2 * c->V * 36, sizeof(float));
   return c;
}

void oc_destroy(oc_ctx *c)
{
   if (!c) return;
   free(c->ipt); free(c->lex); free(c->iup); free(c->idn); free(c->x0);
   free(c->rs); free(c->rd);
   free(c->ud); free(c->u); free(c->swd); free(c->sw); free(c->sw_bgr); free(c->swd_bgr);
   if (c->blk)
   {
      for (int n = 0; n < c->nb; n++)
      {
         free(c->blk[n].ipt); free(c->blk[n].iup); free(c->blk[n].idn);
         free(c->blk[n].imb);
      }
      free(c->blk);
   }
   free(c);
}

int oc_volume(const oc_ctx *c) { return c->V; }
int oc_nthread(const oc_ctx *c) { return c->nthread; }

void oc_get_ipt(const oc_ctx *c, int *ipt)
{
   memcpy(ipt, c->ipt, sizeof(int) * c->V);
}

void oc_get_iupdn(const oc_ctx *c, int *iup, int *idn)
{
   memcpy(iup, c->iup, sizeof(int[4]) * c->V);
   memcpy(idn, c->idn, sizeof(int[4]) * c->V);
}

/* flags/lat_parms.c set_bc_parms: cF' is used only for open-SF bc (type 2) */
void oc_set_bc(oc_ctx *c, int type, double cF0, double cF1)
{
   c->bc = type;
   c->cF[0] = (type < 3) ? cF0 : 1.0;
   c->cF[1] = (type == 2) ? cF1 : c->cF[0];
}

void oc_set_sw(oc_ctx *c, double m0, double csw)
{
   c->m0 = m0;
   c->csw = csw;
}

void oc_set_tm(oc_ctx *c, int eoflg) { c->eoflg = eoflg; }

void oc_set_sap(oc_ctx *c, const int *bs, int isolv, int nmr, int ncy)
{
   memcpy(c->sap_bs, bs, sizeof(int) * 4);
   c->sap_isolv = isolv;
   c->sap_nmr = nmr;
   c->sap_ncy = ncy;
   c->sap_ready = 0;
}

/* ------------------------------------------------------------------ */
/* RANLUX (4 interleaved 48-bit subtract-with-borrow lanes)             */
/* ------------------------------------------------------------------ */

static const int64_t RLX_BASE = (int64_t)0x1000000000000LL;
static const int64_t RLX_MASK = (int64_t)0xffffffffffffLL;

/* seed expansion by a 31-bit shift register, random/ranlux_common.c:73 */
static void rlx_seed(oc_rlx *s, int seed, int flag)
{
   int xbit[31], ibit = 0, jbit = 18;
   for (int k = 0; k < 31; k++)
   {
      xbit[k] = seed & 1;
      seed /= 2;
   }
   for (int lane = 0; lane < 4; lane++)
   {
      for (int k = 0; k < 24; k++)
      {
         int ix = 0;
         for (int l = 0; l < 24; l++)
         {
            ix = 2 * ix + xbit[ibit];
            xbit[ibit] = (xbit[ibit] + xbit[jbit]) % 2;
            ibit = (ibit + 1) % 31;
            jbit = (jbit + 1) % 31;
         }
         int flip = (flag == 0) ? ((k % 4) == lane) : ((k % 4) != lane);
         if (flip) ix = 16777215 - ix;
         if (k & 1)
            s->st[k / 2][lane] += ((int64_t)ix << 24);
         else
            s->st[k / 2][lane] = (int64_t)ix;
      }
   }
   for (int lane = 0; lane < 4; lane++) s->st[12][lane] = 0;
   s->ir = 0;
}

static void rlx_step(oc_rlx *s)
{
   int i = s->ir, j = (s->ir >= 5) ? s->ir - 5 : s->ir + 7;
   for (int k = 0; k < s->pr; k++)
   {
      for (int lane = 0; lane < 4; lane++)
      {
         int64_t d = s->st[j][lane] - s->st[i][lane] - s->st[12][lane];
         s->st[12][lane] = (d < 0);
         s->st[i][lane] = (d + RLX_BASE) & RLX_MASK;
      }
      i = (i + 1) % 12;
      j = (j + 1) % 12;
   }
   s->ir = (s->ir + s->pr) % 12;
}

static void stream_init(oc_stream *st, int dbl, int level, int seed)
{
   memset(st, 0, sizeof(*st));
   st->dbl = dbl;
   if (dbl)
      st->s.pr = (level == 1) ? 202 : 397;
   else
      st->s.pr = (level == 0) ? 109 : ((level == 1) ? 202 : 397);
   rlx_seed(&st->s, seed, dbl);
   st->is = dbl ? 47 : 95;
   st->is_old = 0;
}

static void stream_refill(oc_stream *st)
{
   rlx_step(&st->s);
   if (st->dbl)
   {
      const double onebit = ldexp(1.0, -48);
      for (int k = 0; k < 12; k++)
         for (int lane = 0; lane < 4; lane++)
            st->rd[4 * k + lane] = (double)st->s.st[k][lane] * onebit;
   }
   else
   {
      const float onebit = (float)ldexp(1.0, -24);
      for (int k = 0; k < 12; k++)
         for (int lane = 0; lane < 4; lane++)
         {
            st->rs[8 * k + lane] = (float)(st->s.st[k][lane] & 0xffffff) * onebit;
            st->rs[8 * k + 4 + lane] = (float)(st->s.st[k][lane] >> 24) * onebit;
         }
   }
}

static void stream_draw_d(oc_stream *st, double *r, int n)
{
   for (int k = 0; k < n; k++)
   {
      st->is = (st->is + 1) % 48;
      if (st->is == st->is_old)
      {
         stream_refill(st);
         st->is = 4 * st->s.ir;
         st->is_old = st->is;
      }
      r[k] = st->rd[st->is];
   }
}

static void stream_draw_s(oc_stream *st, float *r, int n)
{
   for (int k = 0; k < n; k++)
   {
      st->is = (st->is + 1) % 96;
      if (st->is == st->is_old)
      {
         stream_refill(st);
         st->is = 8 * st->s.ir;
         st->is_old = st->is;
      }
      r[k] = st->rs[st->is];
   }
}

void oc_start_ranlux(oc_ctx *c, int level, int seed)
{
   /* NPROC=1: loc_seed = seed, thread k seeded with seed + k*max_seed */
   int max_seed = 2147483647 / c->nthread;
   free(c->rs);
   free(c->rd);
   c->rs = malloc(sizeof(oc_stream) * c->nthread);
   c->rd = malloc(sizeof(oc_stream) * c->nthread);
   for (int k = 0; k < c->nthread; k++)
   {
      stream_init(c->rs + k, 0, level, seed + k * max_seed);
      stream_init(c->rd + k, 1, level + 1, seed + k * max_seed);
   }
   c->rlx_init = 1;
}

void oc_ranlxd(oc_ctx *c, int thread, double *r, int n)
{
   stream_draw_d(c->rd + thread, r, n);
}

void oc_ranlxs(oc_ctx *c, int thread, float *r, int n)
{
   stream_draw_s(c->rs + thread, r, n);
}

void oc_kat_ranlxs(int level, int seed, float *r, int n)
{
   oc_stream st;
   stream_init(&st, 0, level, seed);
   stream_draw_s(&st, r, n);
}

void oc_kat_ranlxd(int level, int seed, double *r, int n)
{
   oc_stream st;
   stream_init(&st, 1, level, seed);
   stream_draw_d(&st, r, n);
}

static const double TWOPI = 6.2831853071795865;

void oc_gauss_internal(oc_ctx *c, int thread, double *r, int n)
{
   oc_ranlxd(c, thread, r, n);
   for (int k = 0; k + 1 < n; k += 2)
   {
      double rho = sqrt(-log(1.0 - r[k]));
      double ph = TWOPI * (r[k + 1] - 0.5);
      r[k] = rho * sin(ph);
      r[k + 1] = rho * cos(ph);
   }
   if (n & 1)
   {
      double rho = sqrt(-log(1.0 - r[n - 1])), r1;
      oc_ranlxd(c, thread, &r1, 1);
      r[n - 1] = rho * sin(TWOPI * (r1 - 0.5));
   }
}

void oc_gauss_dble(oc_ctx *c, int thread, double *r, int n)
{
   oc_gauss_internal(c, thread, r, n);
}

static void gauss_s(oc_ctx *c, int thread, float *r, int n)
{
   oc_ranlxs(c, thread, r, n);
   for (int k = 0; k + 1 < n; k += 2)
   {
      double r0 = (double)r[k];
      double r1 = (double)r[k + 1] - 0.5;
      double rho = sqrt(-log(1.0 - r0));
      r1 *= TWOPI;
      r[k] = (float)(rho * sin(r1));
      r[k + 1] = (float)(rho * cos(r1));
   }
}

/* ------------------------------------------------------------------ */
/* gauge field                                                          */
/* ------------------------------------------------------------------ */

/* conj(w x z), include/su3.h:394 */
static void cross_conj(const double *w, const double *z, double *v)
{
   for (int i = 0; i < 3; i++)
   {
      int j = (i + 1) % 3, k = (i + 2) % 3;
      double re = w[2*j] * z[2*k] - w[2*j+1] * z[2*k+1] - (w[2*k] * z[2*j] - w[2*k+1] * z[2*j+1]);
      double im = w[2*j] * z[2*k+1] + w[2*j+1] * z[2*k] - (w[2*k] * z[2*j+1] + w[2*k+1] * z[2*j]);
      v[2*i] = re;
      v[2*i+1] = -im;
   }
}

static void random_vec3(oc_ctx *c, int th, double *v)
{
   double nrm = 0.0;
   while (nrm <= 0.1)
   {
      oc_gauss_internal(c, th, v, 6);
      nrm = 0.0;
      for (int i = 0; i < 6; i++) nrm += v[i] * v[i];
   }
   double f = 1.0 / sqrt(nrm);
   for (int i = 0; i < 6; i++) v[i] *= f;
}

static void random_su3(oc_ctx *c, int th, double *u)
{
   double *r0 = u, *r1 = u + 6, *r2 = u + 12, nrm = 0.0;
   random_vec3(c, th, r0);
   while (nrm <= 0.1)
   {
      random_vec3(c, th, r1);
      cross_conj(r0, r1, r2);
      nrm = 0.0;
      for (int i = 0; i < 6; i++) nrm += r2[i] * r2[i];
   }
   double f = 1.0 / sqrt(nrm);
   for (int i = 0; i < 6; i++) r2[i] *= f;
   cross_conj(r2, r0, r1);
}

void oc_random_ud(oc_ctx *c)
{
   int N0 = c->L[0], h = c->V / 2, bc = c->bc;
   for (int k = 0; k < c->nthread; k++)
   {
      for (int j = k * (c->VT / 2); j < (k + 1) * (c->VT / 2); j++)
      {
         int t = c->x0[j + h];
         double *ud = c->ud + (size_t)18 * 8 * j;
         for (int ifc = 0; ifc < 8; ifc++)
         {
            int draw = 1;
            if (t == 0 && ifc == 1 && bc == 0) draw = 0;
            if (t == 0 && ifc >= 2 && bc == 1) draw = 0;
            if (t == N0 - 1 && ifc == 0 && bc == 0) draw = 0;
            if (draw) random_su3(c, k, ud + 18 * ifc);
         }
      }
   }
   /* open bc: links U(x,0) with x0 = N0-1 vanish (bcnds.c open_bc) */
   if (bc == 0)
   {
      for (int j = 0; j < h; j++)
      {
         int t = c->x0[j + h];
         if (t == 0) memset(c->ud + (size_t)18 * (8 * j + 1), 0, 18 * sizeof(double));
         if (t == N0 - 1) memset(c->ud + (size_t)18 * (8 * j), 0, 18 * sizeof(double));
      }
   }
   c->ud_phase = 0;
}

void oc_set_ud_phase(oc_ctx *c)
{
   /* theta = 0: only the antiperiodic sign of the time links that cross
      the x0 = N0-1 -> 0 boundary (chs_ud0, uflds.c:303-322) */
   if (c->ud_phase) return;
   if (c->bc == 3)
   {
      int N0 = c->L[0], h = c->V / 2;
      for (int j = 0; j < h; j++)
      {
         int t = c->x0[j + h], l = -1;
         if (t == 0) l = 8 * j + 1;
         else if (t == N0 - 1) l = 8 * j;
         if (l >= 0)
            for (int i = 0; i < 18; i++) c->ud[(size_t)18 * l + i] = -c->ud[(size_t)18 * l + i];
      }
   }
   c->ud_phase = 1;
}

double *oc_udfld(oc_ctx *c) { return c->ud; }

void oc_set_ud(oc_ctx *c, const double *ud)
{
   memcpy(c->ud, ud, sizeof(double) * 18 * 4 * (size_t)c->V);
}

void oc_assign_ud2u(oc_ctx *c)
{
   size_t n = (size_t)18 * 4 * c->V;
   for (size_t i = 0; i < n; i++) c->u[i] = (float)c->ud[i];
}

float *oc_ufld(oc_ctx *c) { return c->u; }

/* ------------------------------------------------------------------ */
/* spinor fields                                                        */
/* ------------------------------------------------------------------ */

void oc_random_sd(oc_ctx *c, double *sd, double sigma)
{
   double r[24];
   for (int k = 0; k < c->nthread; k++)
      for (int ix = k * c->VT; ix < (k + 1) * c->VT; ix++)
      {
         oc_gauss_internal(c, k, r, 24);
         for (int i = 0; i < 24; i++) sd[(size_t)24 * ix + i] = sigma * r[i];
      }
}

/* random_sd(vol, icom=0, sd, sigma) at the master thread: stream 0 fills
   vol points from sd (sflds/sflds.c:1254-1268) */
void oc_random_sd_master(oc_ctx *c, int vol, double *sd, double sigma)
{
   double r[24];
   for (int ix = 0; ix < vol; ix++)
   {
      oc_gauss_internal(c, 0, r, 24);
      for (int i = 0; i < 24; i++) sd[(size_t)24 * ix + i] = sigma * r[i];
   }
}

/* random_source of main/ms4.c:383-406: thread k walks the ipt indices
   [k VOLUME_TRD, (k+1) VOLUME_TRD) and draws random_sd(1,0,..) with its
   stream at the points of global time t0, zero elsewhere */
void oc_random_source(oc_ctx *c, int t0, double *sd)
{
   double r[24];
   for (int k = 0; k < c->nthread; k++)
      for (int ix = k * c->VT; ix < (k + 1) * c->VT; ix++)
      {
         if (c->x0[ix] == t0)
         {
            oc_gauss_internal(c, k, r, 24);
            for (int i = 0; i < 24; i++) sd[(size_t)24 * ix + i] = r[i];
         }
         else
            for (int i = 0; i < 24; i++) sd[(size_t)24 * ix + i] = 0.0;
      }
}

void oc_random_s(oc_ctx *c, float *s, float sigma)
{
   float r[24];
   for (int k = 0; k < c->nthread; k++)
      for (int ix = k * c->VT; ix < (k + 1) * c->VT; ix++)
      {
         gauss_s(c, k, r, 24);
         for (int i = 0; i < 24; i++) s[(size_t)24 * ix + i] = sigma * r[i];
      }
}

static int is_bnd_pt(const oc_ctx *c, int ix)
{
   int t = c->x0[ix];
   if (c->bc == 3) return 0;
   return (t == 0) || (t == c->L[0] - 1 && c->bc == 0);
}

void oc_bnd_sd2zero(oc_ctx *c, int set, double *sd)
{
   int h = c->V / 2;
   for (int ix = 0; ix < c->V; ix++)
   {
      int ok = (set == 0) || (set == 1 && ix < h) || (set == 2 && ix >= h);
      if (ok && is_bnd_pt(c, ix)) memset(sd + (size_t)24 * ix, 0, 24 * sizeof(double));
   }
}

void oc_bnd_s2zero(oc_ctx *c, int set, float *s)
{
   int h = c->V / 2;
   for (int ix = 0; ix < c->V; ix++)
   {
      int ok = (set == 0) || (set == 1 && ix < h) || (set == 2 && ix >= h);
      if (ok && is_bnd_pt(c, ix)) memset(s + (size_t)24 * ix, 0, 24 * sizeof(float));
   }
}

double oc_unorm_dble(int vol, const double *sd)
{
   double mx = 0.0;
   for (int ix = 0; ix < vol; ix++)
   {
      double n = 0.0;
      for (int i = 0; i < 24; i++) n += sd[(size_t)24 * ix + i] * sd[(size_t)24 * ix + i];
      if (n > mx) mx = n;
   }
   return sqrt(mx);
}

float oc_unorm(int vol, const float *s)
{
   float mx = 0.0f;
   for (int ix = 0; ix < vol; ix++)
   {
      float n = 0.0f;
      for (int i = 0; i < 24; i++) n += s[(size_t)24 * ix + i] * s[(size_t)24 * ix + i];
      if (n > mx) mx = n;
   }
   return (float)sqrt((double)mx);
}
