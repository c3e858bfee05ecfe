/*
 * oc_flow.c -- Wilson flow of the double-precision gauge field (forward
 * Euler integration).  TEST INFRASTRUCTURE ONLY.
 *
 *   plaquette force  forces/force0.c:211-311 (plaq_frc, plaq_frc_part),
 *                    lattice/uidx.c:237-290 (plaq_uidx, NPROC=1),
 *                    lattice/geogen.c:157-187 (sub-block traversal order)
 *   su3 products     include/su3.h:416-463 (_su3_multiply and the inverse),
 *                    include/su3fcts.h:133-330 (su3xsu3 ... prod2su3alg)
 *   exponential      su3fcts/chexp.c:103-131 (chexp_init), 313-375 (ch2mat),
 *                    377-390 (eval_td), 697-729 (chexp_drv0), 878-929
 *                    (expXsu3)
 *   update           wflow/wflow.c:76-155 (update_ud), 217-249 (fwd_euler)
 *   plaquette sums   uflds/plaq_sum.c:69-191 (plaq_sum_dble, plaq_wsum_dble)
 *
 * The arithmetic is written in the reference's operation order (left to right
 * sums, same association of products), so that with contraction of
 * multiply-adds disabled the results agree with a build of the reference to
 * the last bit.  Links: su3_dble = 18 doubles, row-major (c11.re, c11.im,
 * c12.re, ...); su3_alg_dble = 8 doubles c1..c8; the force field has the
 * layout of the gauge field (4V entries, 8 per odd point).
 */
#include <float.h>
#include <stdlib.h>
#include <string.h>

#include "oc_internal.h"

/* r = u s (3-vectors of complex numbers as 6 doubles) */
static void mulv(const double *u, const double *s, double *r)
{
   for (int i = 0; i < 3; i++)
   {
      const double *a = u + 6 * i;
      r[2 * i] = a[0] * s[0] - a[1] * s[1] + a[2] * s[2] - a[3] * s[3] + a[4] * s[4] - a[5] * s[5];
      r[2 * i + 1] = a[0] * s[1] + a[1] * s[0] + a[2] * s[3] + a[3] * s[2] + a[4] * s[5] + a[5] * s[4];
   }
}

/* r = u^dag s */
static void mulv_dag(const double *u, const double *s, double *r)
{
   for (int i = 0; i < 3; i++)
   {
      const double *a0 = u + 2 * i, *a1 = u + 6 + 2 * i, *a2 = u + 12 + 2 * i;
      r[2 * i] = a0[0] * s[0] + a0[1] * s[1] + a1[0] * s[2] + a1[1] * s[3] + a2[0] * s[4] + a2[1] * s[5];
      r[2 * i + 1] = a0[0] * s[1] - a0[1] * s[0] + a1[0] * s[3] - a1[1] * s[2] + a2[0] * s[5] - a2[1] * s[4];
   }
}

static void get_col(const double *v, int j, double *p)
{
   for (int i = 0; i < 3; i++)
   {
      p[2 * i] = v[6 * i + 2 * j];
      p[2 * i + 1] = v[6 * i + 2 * j + 1];
   }
}

static void put_col(double *w, int j, const double *p)
{
   for (int i = 0; i < 3; i++)
   {
      w[6 * i + 2 * j] = p[2 * i];
      w[6 * i + 2 * j + 1] = p[2 * i + 1];
   }
}

/* w = u v (w may alias v only through the copy below) */
static void su3xsu3(const double *u, const double *v, double *w)
{
   double p[6], q[6], t[18];
   for (int j = 0; j < 3; j++)
   {
      get_col(v, j, p);
      mulv(u, p, q);
      put_col(t, j, q);
   }
   memcpy(w, t, sizeof(t));
}

/* w = u^dag v */
static void su3dagxsu3(const double *u, const double *v, double *w)
{
   double p[6], q[6];
   for (int j = 0; j < 3; j++)
   {
      get_col(v, j, p);
      mulv_dag(u, p, q);
      put_col(w, j, q);
   }
}

/* w = u v^dag: column j of v^dag is the conjugate of row j of v */
static void su3xsu3dag(const double *u, const double *v, double *w)
{
   double p[6], q[6];
   for (int j = 0; j < 3; j++)
   {
      for (int k = 0; k < 3; k++)
      {
         p[2 * k] = v[6 * j + 2 * k];
         p[2 * k + 1] = -v[6 * j + 2 * k + 1];
      }
      mulv(u, p, q);
      put_col(w, j, q);
   }
}

/* X = traceless antihermitian part of w = u v in the su3_alg basis */
static void prod2alg(const double *u, const double *v, double *X)
{
   double p[6], c1[6], c2[6], c3[6];
   get_col(v, 0, p);
   mulv(u, p, c1);
   X[0] = c1[1];
   X[1] = c1[1];
   X[2] = -c1[2];
   X[3] = c1[3];
   X[4] = -c1[4];
   X[5] = c1[5];
   get_col(v, 1, p);
   mulv(u, p, c2);
   X[2] += c2[0];
   X[3] += c2[1];
   X[0] -= c2[3];
   X[6] = -c2[4];
   X[7] = c2[5];
   get_col(v, 2, p);
   mulv(u, p, c3);
   X[4] += c3[0];
   X[5] += c3[1];
   X[6] += c3[2];
   X[7] += c3[3];
   X[1] -= c3[5];
   X[0] *= (1.0 / 3.0);
   X[1] *= (1.0 / 3.0);
   for (int k = 2; k < 8; k++) X[k] *= 0.5;
}

static const int PLNS[6][2] = {{0, 1}, {0, 2}, {0, 3}, {2, 3}, {3, 1}, {1, 2}};

static void frc_part(const oc_ctx *c, double *frc, int ofs, int vol)
{
   const double *ud = c->ud;
   int bc = c->bc, N0 = c->L[0];
   double wd[2][18], X[8];

   for (int ix = ofs; ix < ofs + vol; ix++)
   {
      int t = c->x0[ix];
      if (t < N0 - 1 || bc != 0)
      {
         for (int n = 0; n < 3; n++)
         {
            int mu = PLNS[n][0], nu = PLNS[n][1];
            int ip0 = oc_link_fwd(c, ix, mu), ip1 = oc_link_fwd(c, c->iup[ix][mu], nu);
            int ip2 = oc_link_fwd(c, ix, nu), ip3 = oc_link_fwd(c, c->iup[ix][nu], mu);
            su3xsu3dag(ud + 18 * ip1, ud + 18 * ip3, wd[0]);
            su3dagxsu3(ud + 18 * ip2, ud + 18 * ip0, wd[1]);
            if (t < N0 - 1 || bc == 3)
            {
               prod2alg(wd[0], wd[1], X);
               for (int k = 0; k < 8; k++) frc[8 * ip1 + k] += X[k];
            }
            prod2alg(wd[1], wd[0], X);
            for (int k = 0; k < 8; k++) frc[8 * ip3 + k] -= X[k];
            su3xsu3dag(wd[0], ud + 18 * ip2, wd[1]);
            prod2alg(ud + 18 * ip0, wd[1], X);
            for (int k = 0; k < 8; k++) frc[8 * ip0 + k] += X[k];
            if (t > 0 || bc != 1)
               for (int k = 0; k < 8; k++) frc[8 * ip2 + k] -= X[k];
         }
      }
      if (t > 0 || bc != 1)
      {
         double r = 1.0;
         if ((t == 0 && bc != 3) || (t == N0 - 1 && bc == 0)) r = 0.5;
         for (int n = 3; n < 6; n++)
         {
            int mu = PLNS[n][0], nu = PLNS[n][1];
            int ip0 = oc_link_fwd(c, ix, mu), ip1 = oc_link_fwd(c, c->iup[ix][mu], nu);
            int ip2 = oc_link_fwd(c, ix, nu), ip3 = oc_link_fwd(c, c->iup[ix][nu], mu);
            su3xsu3dag(ud + 18 * ip1, ud + 18 * ip3, wd[0]);
            su3dagxsu3(ud + 18 * ip2, ud + 18 * ip0, wd[1]);
            prod2alg(wd[0], wd[1], X);
            for (int k = 0; k < 8; k++) frc[8 * ip1 + k] += r * X[k];
            prod2alg(wd[1], wd[0], X);
            for (int k = 0; k < 8; k++) frc[8 * ip3 + k] -= r * X[k];
            su3xsu3dag(wd[0], ud + 18 * ip2, wd[1]);
            prod2alg(ud + 18 * ip0, wd[1], X);
            for (int k = 0; k < 8; k++) frc[8 * ip0 + k] += r * X[k];
            for (int k = 0; k < 8; k++) frc[8 * ip2 + k] -= r * X[k];
         }
      }
   }
}

void oc_plaq_frc(oc_ctx *c, double *frc)
{
   int bs[2][4], sbvol[16], sbofs[16];
   for (int mu = 0; mu < 4; mu++)
   {
      bs[0][mu] = c->Ltrd[mu] / 2 + (c->Ltrd[mu] / 2) % 2;
      bs[1][mu] = c->Ltrd[mu] - bs[0][mu];
   }
   for (int ib = 0; ib < 16; ib++)
   {
      int i3 = ib & 1, i2 = (ib >> 1) & 1, i1 = (ib >> 2) & 1, i0 = ib >> 3;
      sbvol[ib] = bs[i0][0] * bs[i1][1] * bs[i2][2] * bs[i3][3];
      sbofs[ib] = 0;
      for (int k = 0; k < ib; k++) sbofs[ib] += sbvol[k];
   }
   memset(frc, 0, sizeof(double) * 8 * 4 * (size_t)c->V);
   for (int isb = 0; isb < 16; isb++)
      for (int k = 0; k < c->nthread; k++)
      {
         int ofs = k * (c->VT / 2) + sbofs[isb] / 2, vol = sbvol[isb] / 2;
         frc_part(c, frc, ofs, vol);
         frc_part(c, frc, ofs + c->V / 2, vol);
      }
}

/* ---- plaquette sums: uflds/plaq_sum.c:69-191 ---- */

/* w = u^dag v^dag: column j is u^dag applied to the conjugate of row j of v
   (su3dagxsu3dag, include/su3fcts.h:235-272) */
static void su3dagxsu3dag(const double *u, const double *v, double *w)
{
   double p[6], q[6];
   for (int j = 0; j < 3; j++)
   {
      for (int k = 0; k < 3; k++)
      {
         p[2 * k] = v[6 * j + 2 * k];
         p[2 * k + 1] = -v[6 * j + 2 * k + 1];
      }
      mulv_dag(u, p, q);
      put_col(w, j, q);
   }
}

/* Re tr(u v), summed as cm3x3_retr (include/su3fcts.h:69-88) */
static double retr(const double *u, const double *v)
{
   double r = 0.0;
   for (int i = 0; i < 3; i++)
      for (int k = 0; k < 3; k++)
      {
         const double *a = u + 6 * i + 2 * k, *b = v + 6 * k + 2 * i;
         double t = a[0] * b[0] - a[1] * b[1];
         if (i == 0 && k == 0)
            r = t;
         else
            r += t;
      }
   return r;
}

/* plaq_dble (plaq_sum.c:69-81) with plaq_uidx (uidx.c:237-290, NPROC=1) */
static double plaq_dble(const oc_ctx *c, int n, int ix)
{
   int mu = PLNS[n][0], nu = PLNS[n][1];
   int ip0 = oc_link_fwd(c, ix, mu), ip1 = oc_link_fwd(c, c->iup[ix][mu], nu);
   int ip2 = oc_link_fwd(c, ix, nu), ip3 = oc_link_fwd(c, c->iup[ix][nu], mu);
   double wd1[18], wd2[18];
   su3xsu3(c->ud + 18 * ip0, c->ud + 18 * ip1, wd1);
   su3dagxsu3dag(c->ud + 18 * ip3, c->ud + 18 * ip2, wd2);
   return retr(wd1, wd2);
}

/* local_plaq_sum_dble (plaq_sum.c:84-152) for boundary types 0 and 3: the
   points of each thread range ix in [k VT, (k+1) VT) in order, one acc_qflt
   per point, then the thread sums in thread order; iw = 1 gives the
   space-like plaquettes on the boundary slices of open bc weight 1/2 */
double oc_plaq_sum(const oc_ctx *c, int iw)
{
   const double wp = iw ? 0.5 : 1.0;
   const int N0 = c->L[0], bc = c->bc;
   double tot[2] = {0.0, 0.0};
   for (int k = 0; k < c->nthread; k++)
   {
      double q[2] = {0.0, 0.0};
      for (int ix = k * c->VT; ix < (k + 1) * c->VT; ix++)
      {
         int t = c->x0[ix];
         double pa = 0.0;
         if (t < N0 - 1 || bc != 0)
            for (int n = 0; n < 3; n++) pa += plaq_dble(c, n, ix);
         if ((t > 0 && t < N0 - 1) || bc == 3)
         {
            for (int n = 3; n < 6; n++) pa += plaq_dble(c, n, ix);
         }
         else
         {
            for (int n = 3; n < 6; n++) pa += wp * plaq_dble(c, n, ix);
         }
         oc_acc_qflt(pa, q);
      }
      oc_add_qflt(tot, q, tot);
   }
   return tot[0];
}

/* ---- SU(3) exponential (Cayley-Hamilton representation) ---- */

static int NCH = 0;
static double CCH[64];

static void chexp_init(void)
{
   if (NCH) return;
   int N = 7;
   double fact = 1.0;
   while (fact > DBL_EPSILON)
   {
      N += 1;
      fact /= (double)(N - 7);
   }
   N += (N % 2);
   CCH[0] = 1.0;
   for (int k = 0; k < N; k++) CCH[k + 1] = CCH[k] / (double)(k + 1);
   NCH = N;
}

/* rows of the matrix X of the algebra element (mapX2v), as 3 complex 3-vectors */
static void alg_rows(const double *X, double v[3][6])
{
   v[0][0] = 0.0;
   v[0][1] = X[0] + X[1];
   v[0][2] = X[2];
   v[0][3] = X[3];
   v[0][4] = X[4];
   v[0][5] = X[5];
   v[1][0] = -X[2];
   v[1][1] = X[3];
   v[1][2] = 0.0;
   v[1][3] = X[1] - 2.0 * X[0];
   v[1][4] = X[6];
   v[1][5] = X[7];
   v[2][0] = -X[4];
   v[2][1] = X[5];
   v[2][2] = -X[6];
   v[2][3] = X[7];
   v[2][4] = 0.0;
   v[2][5] = X[0] - 2.0 * X[1];
}

static double vprod_re(const double *a, const double *b)
{
   return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3] + a[4] * b[4] + a[5] * b[5];
}

static double vprod_im(const double *a, const double *b)
{
   return a[0] * b[1] - a[1] * b[0] + a[2] * b[3] - a[3] * b[2] + a[4] * b[5] - a[5] * b[4];
}

/* u = p0 + p1 X + p2 X^2 (ch2mat) */
static void ch2mat(const double *p, const double *X, double *u)
{
   double v[3][6], zr, zi;
   alg_rows(X, v);
   /* diagonal entries are purely imaginary in v: v[k][2k+1] */
   for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++)
      {
         double *e = u + 6 * i + 2 * j;
         if (i == j)
         {
            e[0] = p[0] - p[3] * v[i][2 * j + 1];
            e[1] = p[1] + p[2] * v[i][2 * j + 1];
         }
         else
         {
            e[0] = p[2] * v[i][2 * j] - p[3] * v[i][2 * j + 1];
            e[1] = p[2] * v[i][2 * j + 1] + p[3] * v[i][2 * j];
         }
      }
   for (int i = 0; i < 3; i++)
   {
      zr = vprod_re(v[i], v[i]);
      u[6 * i + 2 * i] -= p[4] * zr;
      u[6 * i + 2 * i + 1] -= p[5] * zr;
   }
   static const int pr[3][2] = {{0, 1}, {0, 2}, {1, 2}};
   for (int q = 0; q < 3; q++)
   {
      int i = pr[q][0], j = pr[q][1];
      zr = vprod_re(v[i], v[j]);
      zi = vprod_im(v[i], v[j]);
      u[6 * i + 2 * j] -= p[4] * zr + p[5] * zi;
      u[6 * i + 2 * j + 1] -= p[5] * zr - p[4] * zi;
      u[6 * j + 2 * i] -= p[4] * zr - p[5] * zi;
      u[6 * j + 2 * i + 1] -= p[5] * zr + p[4] * zi;
   }
}

/* Cayley-Hamilton coefficients p[3] (complex) of exp(X), |X| <= 1 */
static void chexp_drv0(const double *X, double *p)
{
   double v[4][6], t, d;
   t = 3.0 * (X[0] * X[0] + X[1] * X[1] - X[0] * X[1]) + X[2] * X[2] + X[3] * X[3] + X[4] * X[4] +
       X[5] * X[5] + X[6] * X[6] + X[7] * X[7];
   alg_rows(X, v);
   /* v3 = v1 x v2 (complex cross product, conjugated as _vector_cross_prod) */
   for (int k = 0; k < 3; k++)
   {
      int a = (k + 1) % 3, b = (k + 2) % 3;
      double x1r = v[1][2 * a], x1i = v[1][2 * a + 1], x2r = v[1][2 * b], x2i = v[1][2 * b + 1];
      double y1r = v[2][2 * b], y1i = v[2][2 * b + 1], y2r = v[2][2 * a], y2i = v[2][2 * a + 1];
      v[3][2 * k] = x1r * y1r - x1i * y1i - x2r * y2r + x2i * y2i;
      v[3][2 * k + 1] = x2r * y2i + x2i * y2r - x1r * y1i - x1i * y1r;
   }
   d = vprod_im(v[0], v[3]);
   for (int n = 0; n < 6; n++) p[n] = 0.0;
   p[0] = CCH[NCH - 6];
   for (int n = NCH - 7; n >= 0; n--)
   {
      double q0r = p[0], q0i = p[1], q1r = p[2], q1i = p[3], q2r = p[4], q2i = p[5];
      p[0] = CCH[n] + d * q2i;
      p[1] = -d * q2r;
      p[2] = q0r - t * q2r;
      p[3] = q0i - t * q2i;
      p[4] = q1r;
      p[5] = q1i;
   }
}

/* u <- exp(eps X) u */
void oc_expXsu3(double eps, const double *X, double *u)
{
   double nfrb, Y[8], p[6], ws[2][18];
   int n = 0;
   chexp_init();
   nfrb = 4.0 * (3.0 * (X[0] * X[0] + X[1] * X[1] - X[0] * X[1]) + X[2] * X[2] + X[3] * X[3] +
                 X[4] * X[4] + X[5] * X[5] + X[6] * X[6] + X[7] * X[7]);
   nfrb *= eps * eps;
   while (nfrb > 3.0)
   {
      nfrb *= 0.25;
      eps *= 0.5;
      n += 1;
   }
   for (int k = 0; k < 8; k++) Y[k] = eps * X[k];
   double *u1 = ws[0], *u2 = ws[1];
   chexp_drv0(Y, p);
   ch2mat(p, Y, u2);
   for (int k = 0; k < n; k++)
   {
      double *u3 = u1;
      u1 = u2;
      u2 = u3;
      su3xsu3(u1, u1, u2);
   }
   su3xsu3(u2, u, u);
}

static void update_ud(oc_ctx *c, double eps, const double *frc)
{
   int bc = c->bc, N0 = c->L[0], h = c->V / 2;
   for (int ix = h; ix < c->V; ix++)
   {
      int t = c->x0[ix];
      double *ud = c->ud + 18 * 8 * (size_t)(ix - h);
      const double *fd = frc + 8 * 8 * (size_t)(ix - h);
      for (int ifc = 0; ifc < 8; ifc++)
      {
         double e = eps;
         int upd = 1;
         if (t == 0)
         {
            if (ifc == 1)
               upd = (bc != 0);
            else if (ifc >= 2)
            {
               if (bc == 0 || bc == 2)
                  e = 2.0 * eps;
               else if (bc != 3)
                  upd = 0;
            }
         }
         else if (t == N0 - 1)
         {
            if (ifc == 0)
               upd = (bc != 0);
            else if (ifc >= 2 && bc == 0)
               e = 2.0 * eps;
         }
         if (upd) oc_expXsu3(e, fd + 8 * ifc, ud + 18 * ifc);
      }
   }
}

void oc_fwd_euler(oc_ctx *c, int n, double eps)
{
   if (n <= 0) return;
   double *frc = malloc(sizeof(double) * 8 * 4 * (size_t)c->V);
   for (int k = 0; k < n; k++)
   {
      oc_plaq_frc(c, frc);
      update_ud(c, -eps, frc);
   }
   free(frc);
}
