/*
 * oc_sw.c -- oracle: clover field tensor, SW term, Pauli matrices.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Reference behaviour restated here:
 *   ftensor ................. tcharge/ftensor.c:139-260 (clover of the four
 *                             plaquette leaves, F = (Q - Q^dag)/8, stored as
 *                             u3_alg X_11=i c1, X_12=c4+i c5, ... su3prod.c:56-63)
 *   pauli_term/sw_term ...... sw_term/sw_term.c:154-479.  Upper (lower) 6x6
 *                             block = c2 * sum_k sigma_k (x) i(B_k -/+ E_k),
 *                             B = (F23,F31,F12), E = (F01,F02,F03), c2=-csw/2,
 *                             plus c(x0) on the diagonal; unity on the
 *                             boundary time slices.
 *   pauli packing ........... u[0..5] real diagonal, u[6..35] the 15 upper
 *                             off-diagonal entries row-major as (re,im)
 *                             (sw_term/pauli_dble.c:1336-1440, 1644-1674)
 *   mul_pauli(_dble) ........ pauli_dble.c:1336-1440, pauli.c:953-1057
 *   inv_pauli_dble .......... pauli_dble.c:1718-1758: Hermitian part of
 *                             (A+i mu)^-1 by Householder QR; unsafe if some
 *                             |R_kk|^2 < 1e-4 |M|_F^2 or |M|_F |M^-1|_F > 100
 */
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <complex.h>
#include "oc_internal.h"

typedef double complex zc;

/* ------------------------------------------------------------------ */
/* 3x3 complex helpers (row-major, 9 complex = 18 doubles)              */
/* ------------------------------------------------------------------ */

static void m_load(const double *u, zc *m)
{
   for (int i = 0; i < 9; i++) m[i] = u[2 * i] + I * u[2 * i + 1];
}

static void m_mul(const zc *a, const zc *b, zc *c)
{
   zc t[9];
   for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++)
      {
         zc s = 0;
         for (int k = 0; k < 3; k++) s += a[3 * i + k] * b[3 * k + j];
         t[3 * i + j] = s;
      }
   memcpy(c, t, sizeof(t));
}

static void m_dag(const zc *a, zc *b)
{
   zc t[9];
   for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) t[3 * i + j] = conj(a[3 * j + i]);
   memcpy(b, t, sizeof(t));
}

static void link(oc_ctx *c, int ix, int mu, zc *m)
{
   m_load(c->ud + (size_t)18 * oc_link_fwd(c, ix, mu), m);
}

/* leaf: loop starting at x, first step s1*mu, then s2*nu, back */
static void leaf(oc_ctx *c, int ix, int mu, int nu, int s1, int s2, zc *q)
{
   zc a[9], b[9];
   /* walk the closed path x -> x+s1 mu -> x+s1 mu+s2 nu -> x+s2 nu -> x */
   int y1 = (s1 > 0) ? c->iup[ix][mu] : c->idn[ix][mu];
   int y2 = (s2 > 0) ? c->iup[y1][nu] : c->idn[y1][nu];
   int y3 = (s2 > 0) ? c->iup[ix][nu] : c->idn[ix][nu];
   /* step 1: x -> y1 */
   if (s1 > 0) link(c, ix, mu, q); else { link(c, y1, mu, a); m_dag(a, q); }
   /* step 2: y1 -> y2 */
   if (s2 > 0) link(c, y1, nu, a); else { link(c, y2, nu, b); m_dag(b, a); }
   m_mul(q, a, q);
   /* step 3: y2 -> y3 (direction -s1 mu) */
   if (s1 > 0) { link(c, y3, mu, b); m_dag(b, a); } else link(c, y2, mu, a);
   m_mul(q, a, q);
   /* step 4: y3 -> x (direction -s2 nu) */
   if (s2 > 0) { link(c, ix, nu, b); m_dag(b, a); } else link(c, y3, nu, a);
   m_mul(q, a, q);
}

static const int PLANES[6][2] = {{0, 1}, {0, 2}, {0, 3}, {2, 3}, {3, 1}, {1, 2}};

void oc_ftensor(oc_ctx *c, double *ft)
{
   int N0 = c->L[0];
   for (int n = 0; n < 6; n++)
   {
      int mu = PLANES[n][0], nu = PLANES[n][1];
      for (int ix = 0; ix < c->V; ix++)
      {
         double *f = ft + ((size_t)n * c->V + ix) * 9;
         int t = c->x0[ix];
         memset(f, 0, 9 * sizeof(double));
         if (c->bc == 0 && (t == 0 || t == N0 - 1)) continue;
         zc q[9], l[9];
         memset(q, 0, sizeof(q));
         /* the four leaves, all with the orientation mu -> nu */
         leaf(c, ix, mu, nu, +1, +1, l);
         for (int i = 0; i < 9; i++) q[i] += l[i];
         leaf(c, ix, nu, mu, +1, -1, l);
         for (int i = 0; i < 9; i++) q[i] += l[i];
         leaf(c, ix, mu, nu, -1, -1, l);
         for (int i = 0; i < 9; i++) q[i] += l[i];
         leaf(c, ix, nu, mu, -1, +1, l);
         for (int i = 0; i < 9; i++) q[i] += l[i];
         /* F = (Q - Q^dag)/8 */
         zc F[9];
         for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
               F[3 * i + j] = 0.125 * (q[3 * i + j] - conj(q[3 * j + i]));
         f[0] = cimag(F[0]);
         f[1] = cimag(F[4]);
         f[2] = cimag(F[8]);
         f[3] = creal(F[1]);
         f[4] = cimag(F[1]);
         f[5] = creal(F[2]);
         f[6] = cimag(F[2]);
         f[7] = creal(F[5]);
         f[8] = cimag(F[5]);
      }
   }
}

/* anti-Hermitian 3x3 from u3_alg coefficients */
static void alg_to_mat(const double *x, zc *m)
{
   m[0] = I * x[0];
   m[4] = I * x[1];
   m[8] = I * x[2];
   m[1] = x[3] + I * x[4];
   m[2] = x[5] + I * x[6];
   m[5] = x[7] + I * x[8];
   m[3] = -conj(m[1]);
   m[6] = -conj(m[2]);
   m[7] = -conj(m[5]);
}

static void pack_herm(const zc *H, double diag_shift, double *u)
{
   int p = 0;
   for (int i = 0; i < 6; i++) u[i] = creal(H[6 * i + i]) + diag_shift;
   for (int i = 0; i < 6; i++)
      for (int j = i + 1; j < 6; j++)
      {
         u[6 + 2 * p] = creal(H[6 * i + j]);
         u[7 + 2 * p] = cimag(H[6 * i + j]);
         p++;
      }
}

static void unpack_herm(const double *u, double mu, zc *H)
{
   int p = 0;
   for (int i = 0; i < 6; i++) H[6 * i + i] = u[i] + I * mu;
   for (int i = 0; i < 6; i++)
      for (int j = i + 1; j < 6; j++)
      {
         H[6 * i + j] = u[6 + 2 * p] + I * u[7 + 2 * p];
         H[6 * j + i] = u[6 + 2 * p] - I * u[7 + 2 * p];
         p++;
      }
}

/* sigma_k (x) M into the 6x6 matrix H (spin-major, colour fastest) */
static void add_sigma(int k, const zc *M, zc *H)
{
   static const zc s[3][4] = {{0, 1, 1, 0}, {0, -I, I, 0}, {1, 0, 0, -1}};
   for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++)
      {
         zc sab = s[k][2 * a + b];
         if (sab == 0) continue;
         for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) H[6 * (3 * a + i) + 3 * b + j] += sab * M[3 * i + j];
      }
}

/* Householder-QR inverse of a 6x6 complex matrix; returns the squared
   moduli of the diagonal of R in rkk[] */
static void qr_inverse(const zc *M, zc *Minv, double *rkk)
{
   zc A[36], Q[36];
   memcpy(A, M, sizeof(A));
   for (int i = 0; i < 36; i++) Q[i] = (i % 7 == 0) ? 1.0 : 0.0;
   for (int k = 0; k < 6; k++)
   {
      double nrm = 0.0;
      for (int i = k; i < 6; i++) nrm += creal(A[6 * i + k] * conj(A[6 * i + k]));
      rkk[k] = nrm;
      if (k == 5 || nrm == 0.0) continue;
      double alpha = sqrt(nrm);
      zc a0 = A[6 * k + k];
      zc ph = (cabs(a0) > 0.0) ? a0 / cabs(a0) : 1.0;
      zc v[6];
      for (int i = 0; i < 6; i++) v[i] = (i < k) ? 0 : A[6 * i + k];
      v[k] += ph * alpha;
      double vn = 0.0;
      for (int i = k; i < 6; i++) vn += creal(v[i] * conj(v[i]));
      /* A <- (1 - 2 v v^dag / |v|^2) A, accumulate Q^dag the same way */
      for (int j = 0; j < 6; j++)
      {
         zc s = 0, t = 0;
         for (int i = k; i < 6; i++)
         {
            s += conj(v[i]) * A[6 * i + j];
            t += conj(v[i]) * Q[6 * i + j];
         }
         s *= 2.0 / vn;
         t *= 2.0 / vn;
         for (int i = k; i < 6; i++)
         {
            A[6 * i + j] -= s * v[i];
            Q[6 * i + j] -= t * v[i];
         }
      }
   }
   for (int k = 0; k < 6; k++) rkk[k] = creal(A[6 * k + k] * conj(A[6 * k + k]));
   /* R X = Q^dag  (Q holds Q^dag after the reflections) */
   for (int j = 0; j < 6; j++)
      for (int i = 5; i >= 0; i--)
      {
         zc s = Q[6 * i + j];
         for (int l = i + 1; l < 6; l++) s -= A[6 * i + l] * Minv[6 * l + j];
         Minv[6 * i + j] = (A[6 * i + i] != 0) ? s / A[6 * i + i] : 0;
      }
}

int oc_inv_pauli_dble(double mu, const double *m, double *im)
{
   zc M[36], Mi[36], H[36];
   double rkk[6], nm = 0.0, ni = 0.0;
   unpack_herm(m, mu, M);
   for (int i = 0; i < 36; i++) nm += creal(M[i] * conj(M[i]));
   double eps = 1.0e-4 * nm;
   qr_inverse(M, Mi, rkk);
   int ifail = 0;
   for (int k = 0; k < 6; k++)
      if (rkk[k] < eps) ifail = 1;
   for (int i = 0; i < 36; i++) ni += creal(Mi[i] * conj(Mi[i]));
   if (eps * ni > 1.0) ifail = 1;
   for (int i = 0; i < 6; i++)
      for (int j = 0; j < 6; j++) H[6 * i + j] = 0.5 * (Mi[6 * i + j] + conj(Mi[6 * j + i]));
   pack_herm(H, 0.0, im);
   return ifail;
}

int oc_sw_term(oc_ctx *c, int set)
{
   int N0 = c->L[0], h = c->V / 2, bc = c->bc, ifail = 0;
   double c1 = 4.0 + c->m0, c2 = -0.5 * c->csw;
   double c3[2] = {c1 + c->cF[0] - 1.0, c1 + c->cF[1] - 1.0};
   double *ft = malloc(sizeof(double) * 9 * 6 * (size_t)c->V);
   oc_ftensor(c, ft);
   int inv_e = (set == 0 || set == 1), inv_o = (set == 0 || set == 2);

   for (int ix = 0; ix < c->V; ix++)
   {
      int t = c->x0[ix];
      double *sw = c->swd + (size_t)72 * ix;
      if ((t == 0 && bc != 3) || (t == N0 - 1 && bc == 0))
      {
         memset(sw, 0, 72 * sizeof(double));
         for (int i = 0; i < 6; i++) sw[i] = sw[36 + i] = 1.0;
         continue;
      }
      double cx = c1;
      if (t == 1 && bc != 3) cx = c3[0];
      else if ((t == N0 - 2 && bc == 0) || (t == N0 - 1 && (bc == 1 || bc == 2))) cx = c3[1];
      for (int half = 0; half < 2; half++)
      {
         zc H[36], X[9], A[9], B[9];
         double sg = half ? 1.0 : -1.0;   /* upper: B - E, lower: B + E */
         memset(H, 0, sizeof(H));
         for (int k = 0; k < 3; k++)
         {
            alg_to_mat(ft + ((size_t)(3 + k) * c->V + ix) * 9, B);
            alg_to_mat(ft + ((size_t)k * c->V + ix) * 9, A);
            for (int i = 0; i < 9; i++) X[i] = I * c2 * (B[i] + sg * A[i]);
            add_sigma(k, X, H);
         }
         pack_herm(H, cx, sw + 36 * half);
         if ((ix < h && inv_e) || (ix >= h && inv_o))
            ifail |= oc_inv_pauli_dble(0.0, sw + 36 * half, sw + 36 * half);
      }
   }
   free(ft);
   c->swd_inv[0] = inv_e;
   c->swd_inv[1] = inv_o;
   return ifail;
}

double *oc_swdfld(oc_ctx *c) { return c->swd; }

void oc_assign_swd2sw(oc_ctx *c)
{
   size_t n = (size_t)72 * c->V;
   for (size_t i = 0; i < n; i++) c->sw[i] = (float)c->swd[i];
}

float *oc_swfld(oc_ctx *c) { return c->sw; }

/* r = (A + i mu) s on one Weyl spinor (6 complex) */
void oc_mul_pauli_dble(double mu, const double *m, const double *s, double *r)
{
   zc H[36], v[6], w[6];
   unpack_herm(m, mu, H);
   for (int i = 0; i < 6; i++) v[i] = s[2 * i] + I * s[2 * i + 1];
   for (int i = 0; i < 6; i++)
   {
      zc acc = 0;
      for (int j = 0; j < 6; j++) acc += H[6 * i + j] * v[j];
      w[i] = acc;
   }
   for (int i = 0; i < 6; i++)
   {
      r[2 * i] = creal(w[i]);
      r[2 * i + 1] = cimag(w[i]);
   }
}

/* single precision version: products and sums in float */
void oc_mul_pauli(float mu, const float *m, const float *s, float *r)
{
   float hr[36], hi[36], out[12];
   int p = 0;
   for (int i = 0; i < 6; i++)
   {
      hr[7 * i] = m[i];
      hi[7 * i] = mu;
   }
   for (int i = 0; i < 6; i++)
      for (int j = i + 1; j < 6; j++)
      {
         hr[6 * i + j] = m[6 + 2 * p];
         hi[6 * i + j] = m[7 + 2 * p];
         hr[6 * j + i] = m[6 + 2 * p];
         hi[6 * j + i] = -m[7 + 2 * p];
         p++;
      }
   for (int i = 0; i < 6; i++)
   {
      float re = 0.0f, im = 0.0f;
      for (int j = 0; j < 6; j++)
      {
         re += hr[6 * i + j] * s[2 * j] - hi[6 * i + j] * s[2 * j + 1];
         im += hr[6 * i + j] * s[2 * j + 1] + hi[6 * i + j] * s[2 * j];
      }
      out[2 * i] = re;
      out[2 * i + 1] = im;
   }
   memcpy(r, out, sizeof(out));
}

void oc_apply_sw_dble(int vol, double mu, const double *m, const double *s, double *r)
{
   for (int ix = 0; ix < vol; ix++)
   {
      oc_mul_pauli_dble(mu, m + 72 * (size_t)ix, s + 24 * (size_t)ix, r + 24 * (size_t)ix);
      oc_mul_pauli_dble(-mu, m + 72 * (size_t)ix + 36, s + 24 * (size_t)ix + 12,
                        r + 24 * (size_t)ix + 12);
   }
}
