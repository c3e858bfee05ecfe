/*
 * oc_sap_tmpl.h -- precision-generic block operators, block MinRes and the
 * Schwarz sweep of the CPU oracle (TEST INFRASTRUCTURE ONLY).  Included twice
 * by oc_sap.c with
 *   R      = float            | double
 *   SFX    = (empty)          | _dble        (name suffix)
 *   UFLD   = c->u             | c->ud        (links)
 *   SWB    = c->sw_bgr        | c->swd_bgr   (block SW field, odd part
 *                                             inverted when isolv=1)
 *   MULP   = oc_mul_pauli     | oc_mul_pauli_dble
 *   HOP    = oc_hop_term_f    | oc_hop_term_d
 *   NSQ, SPROD, MULCADD = the BLAS-1 of the precision
 *   RMIN   = FLT_MIN          | DBL_MIN
 *
 * Reference:
 *   Dw_blk, Dwhat_blk, Dwee/Dwoo/Dwoe/Dweo_blk  dirac/Dw.c:1475-1907 (FP32),
 *       dirac/Dw_dble.c:1891-2335 (FP64 twins): Dirichlet exterior (block
 *       neighbour index vol), open-bc boundary points zeroed
 *   blk_mres / blk_eo_mres ....................  sap/blk_solv.c:1117-1186
 *   sap() ......................................  sap/sap.c:592-790
 *   Dw_bnd + sap_com ...........................  dirac/Dw_bnd.c:801-1011,
 *       sap/sap_com.c:465-501 (rho(y) -= hopping term from the block
 *       correction at every exterior neighbour y)
 * The FP64 instance has no reference counterpart for mres and sap (the
 * reference's Schwarz sweep is single precision); it is the same algorithm
 * in double precision, the oracle of the library's FP64 Schwarz variant.
 */

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SFX)
#define ZSR(p, x) memset((p) + 24 * (size_t)(x), 0, 24 * sizeof(R))

static void FN(blk_pauli)(const oc_ctx *c, const oc_block *b, int x, R mu, const R *s, R *r)
{
   const R *m = SWB(c) + (size_t)72 * b->imb[x];
   MULP(mu, m, s, r);
   MULP(-mu, m + 36, s + 12, r + 12);
}

/* r[x] += coef * hopping terms from in-block neighbours of x */
static void FN(blk_hop)(const oc_ctx *c, const oc_block *b, int x, const R *s,
                        const unsigned char *skip, R coef, R *r)
{
   int gx = b->imb[x];
   for (int mu = 0; mu < 4; mu++)
   {
      int y = b->iup[x][mu];
      if (y < b->vol && !(skip && skip[y]))
         HOP(mu, +1, 0, UFLD(c) + (size_t)18 * oc_link_fwd(c, gx, mu), s + 24 * (size_t)y, coef, r);
      y = b->idn[x][mu];
      if (y < b->vol && !(skip && skip[y]))
         HOP(mu, -1, 1, UFLD(c) + (size_t)18 * oc_link_bwd(c, gx, mu), s + 24 * (size_t)y, coef, r);
   }
}

void FN(oc_Dw_blk)(oc_ctx *c, int n, R mu, const R *sin, R *r)
{
   oc_block *b = c->blk + n;
   int vol = b->vol, h = vol / 2;
   R muo = c->eoflg ? (R)0 : mu;
   R *s = malloc(sizeof(R) * 24 * (vol + 1));
   memcpy(s, sin, sizeof(R) * 24 * vol);
   ZSR(s, vol);
   unsigned char *bp = bnd_mask(c, b);
   for (int x = 0; x < vol; x++)
      if (bp[x]) ZSR(s, x);
   for (int x = 0; x < vol; x++)
   {
      R *rr = r + 24 * (size_t)x;
      if (bp[x]) { ZSR(r, x); continue; }
      FN(blk_pauli)(c, b, x, (x < h) ? mu : muo, s + 24 * (size_t)x, rr);
      FN(blk_hop)(c, b, x, s, NULL, (R)-0.5, rr);
   }
   ZSR(r, vol);
   free(bp);
   free(s);
}

/* even-site Schur complement on the block; the odd part of the block SW
   field must be inverted */
void FN(oc_Dwhat_blk)(oc_ctx *c, int n, R mu, const R *sin, R *r)
{
   oc_block *b = c->blk + n;
   int vol = b->vol, h = vol / 2;
   R *s = calloc((size_t)24 * (vol + 1), sizeof(R));
   R *t = calloc((size_t)24 * (vol + 1), sizeof(R));
   memcpy(s, sin, sizeof(R) * 24 * vol);
   unsigned char *bp = bnd_mask(c, b);
   for (int x = 0; x < h; x++) FN(blk_pauli)(c, b, x, mu, s + 24 * (size_t)x, r + 24 * (size_t)x);
   for (int x = 0; x < h; x++)
      if (bp[x]) ZSR(s, x);
   for (int x = h; x < vol; x++)
   {
      if (bp[x]) continue;
      R acc[24] = {0};
      FN(blk_hop)(c, b, x, s, NULL, (R)-0.5, acc);
      FN(blk_pauli)(c, b, x, (R)0, acc, t + 24 * (size_t)x);
   }
   for (int x = 0; x < h; x++) FN(blk_hop)(c, b, x, t, bp, (R)0.5, r + 24 * (size_t)x);
   for (int x = 0; x < h; x++)
      if (bp[x]) ZSR(r, x);
   free(bp);
   free(s);
   free(t);
}

static void FN(blk_Dwoo)(oc_ctx *c, const oc_block *b, R mu, R *s, R *r)
{
   if (c->eoflg) mu = (R)0;
   for (int x = b->vol / 2; x < b->vol; x++)
   {
      if (blk_bnd_pt(c, b, x)) { ZSR(s, x); ZSR(r, x); continue; }
      R out[24];
      FN(blk_pauli)(c, b, x, mu, s + 24 * (size_t)x, out);
      memcpy(r + 24 * (size_t)x, out, sizeof(out));
   }
}

static void FN(blk_Dwoe)(oc_ctx *c, const oc_block *b, R *s, R *r)
{
   int h = b->vol / 2;
   for (int x = 0; x < h; x++)
      if (blk_bnd_pt(c, b, x)) ZSR(s, x);
   for (int x = h; x < b->vol; x++)
   {
      R acc[24] = {0};
      if (!blk_bnd_pt(c, b, x)) FN(blk_hop)(c, b, x, s, NULL, (R)-0.5, acc);
      memcpy(r + 24 * (size_t)x, acc, sizeof(acc));
   }
}

static void FN(blk_Dweo)(oc_ctx *c, const oc_block *b, R *s, R *r)
{
   int h = b->vol / 2;
   unsigned char *bp = bnd_mask(c, b);
   for (int x = h; x < b->vol; x++)
      if (bp[x]) ZSR(s, x);
   for (int x = 0; x < h; x++) FN(blk_hop)(c, b, x, s, bp, (R)0.5, r + 24 * (size_t)x);
   for (int x = 0; x < h; x++)
      if (bp[x]) ZSR(r, x);
   free(bp);
}

static void FN(blk_Dwee)(oc_ctx *c, const oc_block *b, R mu, R *s, R *r)
{
   for (int x = 0; x < b->vol / 2; x++)
   {
      if (blk_bnd_pt(c, b, x)) { ZSR(s, x); ZSR(r, x); continue; }
      R out[24];
      FN(blk_pauli)(c, b, x, mu, s + 24 * (size_t)x, out);
      memcpy(r + 24 * (size_t)x, out, sizeof(out));
   }
}

/* Dwee_blk, Dwoo_blk, Dwoe_blk, Dweo_blk on copies of the block fields
   b.s[k] = s and b.s[l] = r (vol points each, block order); s and r are
   written back, with the zeros the operators put on the boundary points of
   open bc.  op: 0 Dwee, 1 Dwoo, 2 Dwoe, 3 Dweo. */
void FN(oc_blk_eo_op)(oc_ctx *c, int n, int op, R mu, R *s, R *r)
{
   oc_block *b = c->blk + n;
   size_t nb = sizeof(R) * 24 * (size_t)b->vol;
   R *sc = calloc((size_t)24 * (b->vol + 1), sizeof(R));
   R *rc = calloc((size_t)24 * (b->vol + 1), sizeof(R));
   memcpy(sc, s, nb);
   memcpy(rc, r, nb);
   if (op == 0) FN(blk_Dwee)(c, b, mu, sc, rc);
   else if (op == 1) FN(blk_Dwoo)(c, b, mu, sc, rc);
   else if (op == 2) FN(blk_Dwoe)(c, b, sc, rc);
   else FN(blk_Dweo)(c, b, sc, rc);
   memcpy(s, sc, nb);
   memcpy(r, rc, nb);
   free(sc);
   free(rc);
}

/* nmr minimal-residual steps on the block (eo: on the even half with
   Dwhat_blk): s0 = 0; repeat {s2 = D s1; r = |s2|^2; stop if r < 2 RMIN;
   z = (s2, s1) / r; s0 += z s1; s1 -= z s2}  (blk_solv.c:1117-1186) */
static void FN(mres)(oc_ctx *c, int n, R mu, int nmr, int eo, R *s0, R *s1)
{
   int vol = c->blk[n].vol, nv = eo ? vol / 2 : vol;
   R *s2 = calloc((size_t)24 * (vol + 1), sizeof(R));
   memset(s0, 0, sizeof(R) * 24 * nv);
   for (int imr = 0; imr < nmr; imr++)
   {
      if (eo) FN(oc_Dwhat_blk)(c, n, mu, s1, s2);
      else FN(oc_Dw_blk)(c, n, mu, s1, s2);
      R r = NSQ(nv, s2), z[2];
      if (r < (R)2 * RMIN) break;
      SPROD(nv, s2, s1, z);
      r = (R)1 / r;
      z[0] *= r;
      z[1] *= r;
      MULCADD(nv, s0, s1, z[0], z[1]);
      MULCADD(nv, s1, s2, -z[0], -z[1]);
   }
   free(s2);
}

void FN(oc_blk_mres)(oc_ctx *c, int n, R mu, int nmr, R *s0, R *s1)
{
   FN(mres)(c, n, mu, nmr, 0, s0, s1);
}

void FN(oc_blk_eo_mres)(oc_ctx *c, int n, R mu, int nmr, R *s0, R *s1)
{
   FN(mres)(c, n, mu, nmr, 1, s0, s1);
}

/* rho(y) -= hopping term from s0 for every exterior neighbour y */
static void FN(blk_bnd_update)(oc_ctx *c, const oc_block *b, const R *s0, R *rho)
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
            R t[24] = {0};
            HOP(mu, +1, 0, UFLD(c) + (size_t)18 * oc_link_bwd(c, gx, mu), s0 + 24 * (size_t)x, (R)-0.5, t);
            for (int i = 0; i < 24; i++) rho[24 * (size_t)gy + i] -= t[i];
         }
         if (b->iup[x][mu] == b->vol)
         {
            /* y = x + mu receives -1/2 (1+gamma_mu) U(x,mu)^dag s0(x) */
            int gy = c->iup[gx][mu];
            R t[24] = {0};
            HOP(mu, -1, 1, UFLD(c) + (size_t)18 * oc_link_fwd(c, gx, mu), s0 + 24 * (size_t)x, (R)-0.5, t);
            for (int i = 0; i < 24; i++) rho[24 * (size_t)gy + i] -= t[i];
         }
      }
   }
}

void FN(oc_sap)(oc_ctx *c, R mu, int isolv, int nmr, R *psi, R *rho)
{
   if (!c->sap_ready) oc_sap_prepare(c);
   int vol = c->bvol, h = vol / 2;
   R *s0 = calloc((size_t)24 * (vol + 1) * c->nb, sizeof(R));
   R *s1 = calloc((size_t)24 * (vol + 1), sizeof(R));
   for (int ic = 0; ic < 2; ic++)
   {
      int n0 = ic ? c->nbh : 0, n1 = ic ? c->nb : c->nbh;
      for (int n = n0; n < n1; n++)
      {
         oc_block *b = c->blk + n;
         R *b0 = s0 + (size_t)24 * (vol + 1) * n;
         for (int x = 0; x < vol; x++)
            memcpy(s1 + 24 * (size_t)x, rho + 24 * (size_t)b->imb[x], 24 * sizeof(R));
         ZSR(s1, vol);
         if (isolv == 0)
         {
            FN(mres)(c, n, mu, nmr, 0, b0, s1);
            for (int x = 0; x < vol; x++)
            {
               R *p = psi + 24 * (size_t)b->imb[x], *q = rho + 24 * (size_t)b->imb[x];
               for (int i = 0; i < 24; i++) p[i] += b0[24 * (size_t)x + i];
               memcpy(q, s1 + 24 * (size_t)x, 24 * sizeof(R));
            }
         }
         else
         {
            FN(blk_Dwoo)(c, b, (R)0, s1, s1);
            FN(blk_Dweo)(c, b, s1, s1);
            FN(mres)(c, n, mu, nmr, 1, b0, s1);
            FN(blk_Dwoe)(c, b, b0, b0);
            FN(blk_Dwoo)(c, b, (R)0, b0, b0);
            for (int x = 0; x < vol; x++)
            {
               R *p = psi + 24 * (size_t)b->imb[x], *q = rho + 24 * (size_t)b->imb[x];
               R *a = b0 + 24 * (size_t)x, *r1 = s1 + 24 * (size_t)x;
               if (x < h)
               {
                  for (int i = 0; i < 24; i++) p[i] += a[i];
                  memcpy(q, r1, 24 * sizeof(R));
               }
               else
               {
                  for (int i = 0; i < 24; i++)
                  {
                     a[i] = r1[i] - a[i];
                     p[i] += a[i];
                  }
                  if (c->eoflg)
                     memset(q, 0, 24 * sizeof(R));
                  else
                     for (int i = 0; i < 12; i++)
                     {
                        /* rho = -i mu gamma5 s0: spins 1,2 get -i mu, spins 3,4 +i mu */
                        R cm = (i < 6) ? -mu : mu;
                        q[2 * i] = -cm * a[2 * i + 1];
                        q[2 * i + 1] = cm * a[2 * i];
                     }
               }
            }
         }
      }
      for (int n = n0; n < n1; n++)
         FN(blk_bnd_update)(c, c->blk + n, s0 + (size_t)24 * (vol + 1) * n, rho);
   }
   free(s0);
   free(s1);
}

#undef ZSR
#undef FN
#undef CAT
#undef CAT2
