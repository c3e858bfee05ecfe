/*
 * oc_dirac_tmpl.h -- precision-generic Wilson-clover operator for the CPU
 * oracle (TEST INFRASTRUCTURE ONLY).  Included twice by oc_dirac.c with
 *   R    = double | float        (arithmetic type)
 *   SFX  = _dble  | (empty)       (name suffix)
 *   UFLD / SWFLD = c->ud, c->swd | c->u, c->sw
 *
 * The operator is written in gather form: every output site collects the
 * eight hopping terms
 *     -1/2 [ (1-gamma_mu) U(x,mu) s(x+mu) + (1+gamma_mu) U(x-mu,mu)^dag s(x-mu) ]
 * and the Pauli term.  The boundary-time rules (which sites are zeroed in
 * input and output, which hopping terms are skipped) follow
 * dirac/Dw_dble.c:1377-1888 and dirac/Dw.c:973-1474 exactly; they are the
 * same in both precisions.  At NPROC=1 the halo routines cps(d)_int_bnd and
 * cps(d)_ext_bnd called by Dw, Dwoe, Dweo and Dwhat zero the even boundary
 * points of the input (int) and of the output (ext), sflds/sdcom.c:593,504
 * and sflds/scom.c:485,573.
 */

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SFX)

/* spin structure of gamma_mu (devel/dirac/check5.c:52-103):
   (gamma_mu s)_a = ph[mu][a] * s_perm[mu][a] with ph in {+-1, +-i} encoded
   as (re, im) pairs */
static const int FN(g_perm)[4][4] = {{2, 3, 0, 1}, {3, 2, 1, 0}, {3, 2, 1, 0}, {2, 3, 0, 1}};
static const int FN(g_ph)[4][4][2] = {
   {{-1, 0}, {-1, 0}, {-1, 0}, {-1, 0}},
   {{0, -1}, {0, -1}, {0, 1}, {0, 1}},
   {{-1, 0}, {1, 0}, {1, 0}, {-1, 0}},
   {{0, -1}, {0, 1}, {0, 1}, {0, -1}}};

/* r (4 spin x 3 colour complex) += coef * (1 - sgn*gamma_mu) (M v) where M is
   U or U^dag */
static void FN(hop_term)(int mu, int sgn, int dag, const R *u, const R *v, R coef, R *r)
{
   R t[24];
   for (int a = 0; a < 4; a++)
      for (int i = 0; i < 3; i++)
      {
         R re = 0, im = 0;
         for (int j = 0; j < 3; j++)
         {
            R ur, ui;
            if (dag) { ur = u[2 * (3 * j + i)]; ui = -u[2 * (3 * j + i) + 1]; }
            else     { ur = u[2 * (3 * i + j)]; ui = u[2 * (3 * i + j) + 1]; }
            R vr = v[6 * a + 2 * j], vi = v[6 * a + 2 * j + 1];
            re += ur * vr - ui * vi;
            im += ur * vi + ui * vr;
         }
         t[6 * a + 2 * i] = re;
         t[6 * a + 2 * i + 1] = im;
      }
   for (int a = 0; a < 4; a++)
   {
      int b = FN(g_perm)[mu][a];
      R pr = (R)FN(g_ph)[mu][a][0], pi = (R)FN(g_ph)[mu][a][1];
      for (int i = 0; i < 3; i++)
      {
         R tr = t[6 * b + 2 * i], ti = t[6 * b + 2 * i + 1];
         /* (gamma t)_a = ph * t_b */
         R gr = pr * tr - pi * ti, gi = pr * ti + pi * tr;
         r[6 * a + 2 * i] += coef * (t[6 * a + 2 * i] - (R)sgn * gr);
         r[6 * a + 2 * i + 1] += coef * (t[6 * a + 2 * i + 1] - (R)sgn * gi);
      }
   }
}

/* r += coef * sum_mu hopping terms at site ix from field s; neighbours whose
   mask entry is set contribute zero */
static void FN(hop_site)(oc_ctx *c, int ix, const R *s, const unsigned char *zero, R coef, R *r)
{
   const R *U = UFLD(c);
   for (int mu = 0; mu < 4; mu++)
   {
      int iy = c->iup[ix][mu];
      if (!zero || !zero[iy])
         FN(hop_term)(mu, +1, 0, U + (size_t)18 * oc_link_fwd(c, ix, mu), s + (size_t)24 * iy, coef, r);
      iy = c->idn[ix][mu];
      if (!zero || !zero[iy])
         FN(hop_term)(mu, -1, 1, U + (size_t)18 * oc_link_bwd(c, ix, mu), s + (size_t)24 * iy, coef, r);
   }
}

static void FN(pauli_site)(oc_ctx *c, int ix, R mu, const R *s, R *r)
{
   const R *m = SWFLD(c) + (size_t)72 * ix;
   MULP(mu, m, s, r);
   MULP(-mu, m + 36, s + 12, r + 12);
}

static int FN(bnd_site)(const oc_ctx *c, int ix)
{
   int t = c->x0[ix];
   return (t == 0 && c->bc != 3) || (t == c->L[0] - 1 && c->bc == 0);
}

static unsigned char *FN(odd_bnd_mask)(oc_ctx *c)
{
   unsigned char *z = calloc(c->V, 1);
   for (int ix = c->V / 2; ix < c->V; ix++) z[ix] = (unsigned char)FN(bnd_site)(c, ix);
   return z;
}

/* cps(d)_int_bnd / cps(d)_ext_bnd at NPROC=1: zero the even boundary points */
static void FN(even_bnd_zero)(oc_ctx *c, R *f)
{
   for (int ix = 0; ix < c->V / 2; ix++)
      if (FN(bnd_site)(c, ix)) memset(f + (size_t)24 * ix, 0, 24 * sizeof(R));
}

/* Dw: even sites  r = A s + hop(odd s, odd boundary sites zero);
       odd interior r = A s (mu or 0 if eoflg) + hop(even s);
       odd boundary sites: s = r = 0  (Dw_dble.c:1377-1504) */
void CAT(oc_Dw, SFX)(oc_ctx *c, R mu, R *s, R *r)
{
   int h = c->V / 2;
   R muo = c->eoflg ? (R)0 : mu;
   unsigned char *z = FN(odd_bnd_mask)(c);
   FN(even_bnd_zero)(c, s);
   for (int ix = h; ix < c->V; ix++)
      if (z[ix]) memset(s + (size_t)24 * ix, 0, 24 * sizeof(R));
   for (int ix = 0; ix < c->V; ix++)
   {
      R *rr = r + (size_t)24 * ix;
      if (ix >= h && z[ix])
      {
         memset(rr, 0, 24 * sizeof(R));
         continue;
      }
      FN(pauli_site)(c, ix, (ix < h) ? mu : muo, s + (size_t)24 * ix, rr);
      FN(hop_site)(c, ix, s, (ix < h) ? z : NULL, (R)-0.5, rr);
   }
   FN(even_bnd_zero)(c, r);
   free(z);
}

void CAT(oc_Dwee, SFX)(oc_ctx *c, R mu, R *s, R *r)
{
   for (int ix = 0; ix < c->V / 2; ix++)
   {
      if (FN(bnd_site)(c, ix))
      {
         memset(s + (size_t)24 * ix, 0, 24 * sizeof(R));
         memset(r + (size_t)24 * ix, 0, 24 * sizeof(R));
      }
      else
         FN(pauli_site)(c, ix, mu, s + (size_t)24 * ix, r + (size_t)24 * ix);
   }
}

void CAT(oc_Dwoo, SFX)(oc_ctx *c, R mu, R *s, R *r)
{
   if (c->eoflg) mu = 0;
   for (int ix = c->V / 2; ix < c->V; ix++)
   {
      if (FN(bnd_site)(c, ix))
      {
         memset(s + (size_t)24 * ix, 0, 24 * sizeof(R));
         memset(r + (size_t)24 * ix, 0, 24 * sizeof(R));
      }
      else
         FN(pauli_site)(c, ix, mu, s + (size_t)24 * ix, r + (size_t)24 * ix);
   }
}

/* r_odd = -1/2 hop(s_even); zero on odd boundary sites */
void CAT(oc_Dwoe, SFX)(oc_ctx *c, R *s, R *r)
{
   FN(even_bnd_zero)(c, s);
   for (int ix = c->V / 2; ix < c->V; ix++)
   {
      R *rr = r + (size_t)24 * ix;
      R acc[24] = {0};
      if (!FN(bnd_site)(c, ix)) FN(hop_site)(c, ix, s, NULL, (R)-0.5, acc);
      memcpy(rr, acc, sizeof(acc));
   }
}

/* r_even -= -1/2 hop(s_odd); odd boundary input sites are zeroed first */
void CAT(oc_Dweo, SFX)(oc_ctx *c, R *s, R *r)
{
   unsigned char *z = FN(odd_bnd_mask)(c);
   for (int ix = c->V / 2; ix < c->V; ix++)
      if (z[ix]) memset(s + (size_t)24 * ix, 0, 24 * sizeof(R));
   for (int ix = 0; ix < c->V / 2; ix++)
      FN(hop_site)(c, ix, s, z, (R)0.5, r + (size_t)24 * ix);
   FN(even_bnd_zero)(c, r);
   free(z);
}

/* r_even = A_ee s_e (mu) + 1/4 hop( A_oo^-1 hop(s_e) ) with the odd boundary
   sites left out (Dw_dble.c:1792-1888); A_oo must hold the inverse */
void CAT(oc_Dwhat, SFX)(oc_ctx *c, R mu, R *s, R *r)
{
   int h = c->V / 2;
   R *t = calloc((size_t)24 * c->V, sizeof(R));
   unsigned char *z = FN(odd_bnd_mask)(c);
   FN(even_bnd_zero)(c, s);
   for (int ix = h; ix < c->V; ix++)
   {
      if (z[ix]) continue;
      R acc[24] = {0};
      FN(hop_site)(c, ix, s, NULL, (R)-0.5, acc);
      FN(pauli_site)(c, ix, (R)0, acc, t + (size_t)24 * ix);
   }
   for (int ix = 0; ix < h; ix++)
   {
      R *rr = r + (size_t)24 * ix;
      FN(pauli_site)(c, ix, mu, s + (size_t)24 * ix, rr);
      FN(hop_site)(c, ix, t, z, (R)0.5, rr);
   }
   FN(even_bnd_zero)(c, r);
   free(z);
   free(t);
}

#undef FN
#undef CAT
#undef CAT2
