/*
 * oc_linalg.c -- oracle: spinor BLAS-1 and double-double sums.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 *   acc_qflt / add_qflt ...... utils/qsum.c:265-310 (two-sum accumulation of
 *                              a double into a (hi,lo) pair; pair + pair)
 *   spinor_prod_dble etc. .... linalg/salg_dble.c:2319-2450: plain double sums
 *                              over chunks of 8 sites, each chunk folded into
 *                              a double-double accumulator
 *   spinor_prod / norm_square  linalg/salg.c:1050-1120, 1223-1330: products in
 *                              float, per-site sums cast to double and summed
 *                              in double, result cast to float
 *   mulc_spinor_add .......... salg.c:1123-1137, salg_dble.c
 */
#include <math.h>
#include <stddef.h>
#include "oc_internal.h"

void oc_acc_qflt(double u, double *qr)
{
   double a = qr[0] + u;
   double qp = a - u;
   double up = a - qp;
   double b = (qr[0] - qp) + (u - up);
   double c = qr[1] + b;
   double d = a + c;
   qr[0] = d;
   qr[1] = c - (d - a);
}

void oc_add_qflt(const double *qu, const double *qv, double *qr)
{
   double a = qu[0] + qv[0];
   double c = qu[1] + qv[1];
   double up = a - qv[0];
   double rp = c - qv[1];
   double vp = a - up;
   double sp = c - rp;
   double b = (qu[0] - up) + (qv[0] - vp);
   double d = (qu[1] - rp) + (qv[1] - sp);
   b += c;
   double e = a + b;
   double f = b - (e - a);
   f += d;
   double w = e + f;
   qr[0] = w;
   qr[1] = f - (w - e);
}

/* Re and Im of (r,s) over a whole spinor: the reference's macros expand
   to one left-to-right sum of 24 products (salg_dble.c:2340-2350) */
static double sre_d(const double *r, const double *s)
{
   double t = r[0] * s[0];
   for (int i = 1; i < 24; i++) t = t + r[i] * s[i];
   return t;
}

static double sim_d(const double *r, const double *s)
{
   double t = r[0] * s[1];
   t = t - r[1] * s[0];
   for (int i = 2; i < 24; i += 2)
   {
      t = t + r[i] * s[i + 1];
      t = t - r[i + 1] * s[i];
   }
   return t;
}

static float vre_f(const float *r, const float *s)
{
   return r[0] * s[0] + r[1] * s[1] + r[2] * s[2] + r[3] * s[3] + r[4] * s[4] + r[5] * s[5];
}

static float vim_f(const float *r, const float *s)
{
   return r[0] * s[1] - r[1] * s[0] + r[2] * s[3] - r[3] * s[2] + r[4] * s[5] - r[5] * s[4];
}

void oc_spinor_prod_dble(int vol, const double *s, const double *r, double *qre, double *qim)
{
   qre[0] = qre[1] = qim[0] = qim[1] = 0.0;
   for (int ix0 = 0; ix0 < vol; ix0 += 8)
   {
      int ix1 = (ix0 + 8 < vol) ? ix0 + 8 : vol;
      double zr = 0.0, zi = 0.0;
      for (int ix = ix0; ix < ix1; ix++)
      {
         const double *a = s + 24 * (size_t)ix, *b = r + 24 * (size_t)ix;
         zr += sre_d(a, b);
         zi += sim_d(a, b);
      }
      oc_acc_qflt(zr, qre);
      oc_acc_qflt(zi, qim);
   }
}

void oc_norm_square_dble(int vol, const double *s, double *q)
{
   q[0] = q[1] = 0.0;
   for (int ix0 = 0; ix0 < vol; ix0 += 8)
   {
      int ix1 = (ix0 + 8 < vol) ? ix0 + 8 : vol;
      double x = 0.0;
      for (int ix = ix0; ix < ix1; ix++)
      {
         const double *a = s + 24 * (size_t)ix;
         x += sre_d(a, a);
      }
      oc_acc_qflt(x, q);
   }
}

void oc_spinor_prod(int vol, const float *s, const float *r, float *z)
{
   double vr = 0.0, vi = 0.0;
   for (int ix = 0; ix < vol; ix++)
   {
      const float *a = s + 24 * (size_t)ix, *b = r + 24 * (size_t)ix;
      vr += (double)vre_f(a, b) + (double)vre_f(a + 6, b + 6) + (double)vre_f(a + 12, b + 12) +
            (double)vre_f(a + 18, b + 18);
      vi += (double)vim_f(a, b) + (double)vim_f(a + 6, b + 6) + (double)vim_f(a + 12, b + 12) +
            (double)vim_f(a + 18, b + 18);
   }
   z[0] = (float)vr;
   z[1] = (float)vi;
}

float oc_norm_square(int vol, const float *s)
{
   double x = 0.0;
   for (int ix = 0; ix < vol; ix++)
   {
      const float *a = s + 24 * (size_t)ix;
      x += (double)vre_f(a, a) + (double)vre_f(a + 6, a + 6) + (double)vre_f(a + 12, a + 12) +
           (double)vre_f(a + 18, a + 18);
   }
   return (float)x;
}

void oc_mulc_spinor_add_dble(int vol, double *s, const double *r, double zre, double zim)
{
   for (size_t i = 0; i < (size_t)12 * vol; i++)
   {
      double rr = r[2 * i], ri = r[2 * i + 1];
      s[2 * i] += zre * rr - zim * ri;
      s[2 * i + 1] += zre * ri + zim * rr;
   }
}

void oc_mulc_spinor_add(int vol, float *s, const float *r, float zre, float zim)
{
   for (size_t i = 0; i < (size_t)12 * vol; i++)
   {
      float rr = r[2 * i], ri = r[2 * i + 1];
      s[2 * i] += zre * rr - zim * ri;
      s[2 * i + 1] += zre * ri + zim * rr;
   }
}
