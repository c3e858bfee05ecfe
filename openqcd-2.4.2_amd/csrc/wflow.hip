// wflow.hip -- Wilson flow of the double-precision gauge field on gfx950.
//
// Replaces wflow/wflow.c:76-249 (update_ud, fwd_euler), forces/force0.c:
// 211-311 (plaq_frc) and su3fcts/chexp.c:103-929 (the SU(3) exponential), and
// provides set_ud_phase/unset_ud_phase of uflds/uflds.c:338-400 for the
// configurations used here (theta = 0).
//
// The reference visits the points x and planes (mu,nu) and adds the
// traceless antihermitian part of each plaquette, started at each of its
// four links, to the force on that link.  Here one lane per link gathers the
// six plaquettes that contain its link (two per plane), computing every term
// with the same products and the same association as the reference:
//
//   plaquette at b in plane (mu,nu):  U0=U(b,mu) U1=U(b+mu,nu) U2=U(b,nu)
//   U3=U(b+nu,mu);  F(U0) += P(U0 (U1 U3^+) U2^+),  F(U2) -= the same,
//   F(U1) += P((U1 U3^+)(U2^+ U0)),  F(U3) -= P((U2^+ U0)(U1 U3^+))
//
// with P(w) = (w - w^+)/2 - tr(w - w^+)/6 expressed in the 8 real components
// of su3_alg_dble.  Only the order in which the six terms of a link are added
// differs from the reference, so forces agree to rounding.  Open boundary
// conditions (type 0): time-like plaquettes at x0 = N0-1 are omitted and the
// space-like plaquettes on the boundary slices carry weight 1/2
// (force0.c:226-274); the update multiplies the step of the space-like links
// on the boundary slices by 2 and leaves the time-like links at x0 = N0-1
// fixed (wflow.c:96-144).
//
// Force field: 8 reals per link stored as 4 double2 units, 32 units per odd
// point in the site-block layout of the gauge field (link l: units 4l..4l+3).
#include <cfloat>
#include <cmath>

#include "dirac_dev.h"
#include "ddsum.h"

namespace oqh {

namespace {

struct Su3
{
   double2 e[9];
};

// all products are evaluated without contraction, in the order of the
// reference's macros (include/su3.h:416-463), so that every plaquette term
// is the reference's value
#pragma clang fp contract(off)

__device__ inline double2 dot_row(const double2 *a, const double2 *s)
{
   // sum_k a_k s_k, left to right as _su3_multiply
   double re = a[0].x * s[0].x - a[0].y * s[0].y + a[1].x * s[1].x - a[1].y * s[1].y + a[2].x * s[2].x - a[2].y * s[2].y;
   double im = a[0].x * s[0].y + a[0].y * s[0].x + a[1].x * s[1].y + a[1].y * s[1].x + a[2].x * s[2].y + a[2].y * s[2].x;
   return make_double2(re, im);
}

__device__ inline double2 dot_col_dag(const double2 *u, int i, const double2 *s)
{
   // (u^+ s)_i, as _su3_inverse_multiply
   const double2 a0 = u[i], a1 = u[3 + i], a2 = u[6 + i];
   double re = a0.x * s[0].x + a0.y * s[0].y + a1.x * s[1].x + a1.y * s[1].y + a2.x * s[2].x + a2.y * s[2].y;
   double im = a0.x * s[0].y - a0.y * s[0].x + a1.x * s[1].y - a1.y * s[1].x + a2.x * s[2].y - a2.y * s[2].x;
   return make_double2(re, im);
}

// w = u v
__device__ inline Su3 mul(const Su3 &u, const Su3 &v)
{
   Su3 w;
#pragma unroll
   for (int j = 0; j < 3; j++)
   {
      double2 p[3] = {v.e[j], v.e[3 + j], v.e[6 + j]};
#pragma unroll
      for (int i = 0; i < 3; i++) w.e[3 * i + j] = dot_row(u.e + 3 * i, p);
   }
   return w;
}

// w = u^+ v
__device__ inline Su3 dag_mul(const Su3 &u, const Su3 &v)
{
   Su3 w;
#pragma unroll
   for (int j = 0; j < 3; j++)
   {
      double2 p[3] = {v.e[j], v.e[3 + j], v.e[6 + j]};
#pragma unroll
      for (int i = 0; i < 3; i++) w.e[3 * i + j] = dot_col_dag(u.e, i, p);
   }
   return w;
}

// w = u v^+
__device__ inline Su3 mul_dag(const Su3 &u, const Su3 &v)
{
   Su3 w;
#pragma unroll
   for (int j = 0; j < 3; j++)
   {
      double2 p[3];
#pragma unroll
      for (int k = 0; k < 3; k++) p[k] = make_double2(v.e[3 * j + k].x, -v.e[3 * j + k].y);
#pragma unroll
      for (int i = 0; i < 3; i++) w.e[3 * i + j] = dot_row(u.e + 3 * i, p);
   }
   return w;
}

// X = P(u v) in the su3_alg_dble basis (include/su3fcts.h:275-330)
__device__ inline void prod2alg(const Su3 &u, const Su3 &v, double *X)
{
   double2 p[3], c1[3], c2[3], c3[3];
#pragma unroll
   for (int j = 0; j < 3; j++)
   {
      p[0] = v.e[j];
      p[1] = v.e[3 + j];
      p[2] = v.e[6 + j];
      double2 *c = j == 0 ? c1 : (j == 1 ? c2 : c3);
#pragma unroll
      for (int i = 0; i < 3; i++) c[i] = dot_row(u.e + 3 * i, p);
   }
   X[0] = c1[0].y;
   X[1] = c1[0].y;
   X[2] = -c1[1].x;
   X[3] = c1[1].y;
   X[4] = -c1[2].x;
   X[5] = c1[2].y;
   X[2] += c2[0].x;
   X[3] += c2[0].y;
   X[0] -= c2[1].y;
   X[6] = -c2[2].x;
   X[7] = c2[2].y;
   X[4] += c3[0].x;
   X[5] += c3[0].y;
   X[6] += c3[1].x;
   X[7] += c3[1].y;
   X[1] -= c3[2].y;
   X[0] *= (1.0 / 3.0);
   X[1] *= (1.0 / 3.0);
#pragma unroll
   for (int k = 2; k < 8; k++) X[k] *= 0.5;
}

struct FArgs
{
   const double2 *ud;
   double2 *u_w;     // gauge field written by the update
   double2 *frc;     // 32 units per odd point
   const int *nbr;   // [8][V], bit 31 flags boundary points
   const int *x0;
   int V, h, N0, bc;
   double eps;
};

// U(z, d) for any point z: odd z stores it in slot 2d, even z in slot 2d+1
// of its forward neighbour
__device__ inline Su3 link_at(const FArgs &a, int z, int d)
{
   int o, l;
   if (z >= a.h)
   {
      o = z;
      l = 2 * d;
   }
   else
   {
      o = a.nbr[(size_t)(2 * d) * a.V + z] & 0x7fffffff;
      l = 2 * d + 1;
   }
   const double2 *p = link_ptr(a.ud, o - a.h, l);
   Su3 u;
#pragma unroll
   for (int k = 0; k < 9; k++) u.e[k] = p[k * SB];
   return u;
}

__device__ inline int up(const FArgs &a, int z, int d) { return a.nbr[(size_t)(2 * d) * a.V + z] & 0x7fffffff; }
__device__ inline int dn(const FArgs &a, int z, int d) { return a.nbr[(size_t)(2 * d + 1) * a.V + z] & 0x7fffffff; }

__device__ inline void acc8(double *F, const double *X, double r)
{
#pragma unroll
   for (int k = 0; k < 8; k++) F[k] += r * X[k];
}

// the planes (mu,nu) of force0.c / uidx.c:70 in their orientation
__device__ inline int plane_n(int a, int b)
{
   // returns n with {PLNS[n]} = {a,b}; PLNS = (0,1),(0,2),(0,3),(2,3),(3,1),(1,2)
   int lo = a < b ? a : b, hi = a < b ? b : a;
   if (lo == 0) return hi - 1;
   if (lo == 2) return 3;            // (2,3)
   if (hi == 3) return 4;            // (1,3) -> (3,1)
   return 5;                         // (1,2)
}

__device__ inline void plane_pq(int n, int &p, int &q)
{
   const int P[6] = {0, 0, 0, 2, 3, 1}, Q[6] = {1, 2, 3, 3, 1, 2};
   p = P[n];
   q = Q[n];
}

// weights of a plaquette with base point time t (force0.c:221-274); returns
// false if the plaquette does not contribute; *w1 tells whether the term of
// its link U1 is taken
__device__ inline bool plaq_weight(const FArgs &a, int n, int t, double *r, bool *w1)
{
   *w1 = true;
   *r = 1.0;
   if (n < 3)
   {
      if (!(t < a.N0 - 1 || a.bc != 0)) return false;
      *w1 = (t < a.N0 - 1 || a.bc == 3);
      return true;
   }
   if ((t == 0 && a.bc != 3) || (t == a.N0 - 1 && a.bc == 0)) *r = 0.5;
   return true;
}

}  // namespace

// one lane per link: link (odd point o, slot l); its base point is o for
// even l and o - mu for odd l
__global__ __launch_bounds__(256) void k_plaq_frc(FArgs a)
{
   int k = blockIdx.x * 256 + threadIdx.x;
   if (k >= 4 * a.V) return;
   int j = k >> 3, l = k & 7;
   int o = a.h + j, mu = l >> 1;
   int y = (l & 1) ? dn(a, o, mu) : o;
   double F[8] = {0, 0, 0, 0, 0, 0, 0, 0}, X[8];
   Su3 Ul = link_at(a, y, mu);
   for (int nu = 0; nu < 4; nu++)
   {
      if (nu == mu) continue;
      int n = plane_n(mu, nu), p, q;
      plane_pq(n, p, q);
      double r;
      bool w1;
      if (mu == p)
      {
         // (i) base y: the link is U0
         int b = y;
         if (plaq_weight(a, n, a.x0[b], &r, &w1))
         {
            Su3 U1 = link_at(a, up(a, b, p), q), U2 = link_at(a, b, q), U3 = link_at(a, up(a, b, q), p);
            Su3 w0 = mul_dag(U1, U3);
            Su3 w2 = mul_dag(w0, U2);
            prod2alg(Ul, w2, X);
            acc8(F, X, r);
         }
         // (ii) base y - q: the link is U3
         b = dn(a, y, q);
         if (plaq_weight(a, n, a.x0[b], &r, &w1))
         {
            Su3 U0 = link_at(a, b, p), U1 = link_at(a, up(a, b, p), q), U2 = link_at(a, b, q);
            Su3 w0 = mul_dag(U1, Ul);
            Su3 wb = dag_mul(U2, U0);
            prod2alg(wb, w0, X);
            acc8(F, X, -r);
         }
      }
      else
      {
         // (iii) base y: the link is U2 (terms with t = 0 of SF bc excluded)
         int b = y;
         if (plaq_weight(a, n, a.x0[b], &r, &w1))
         {
            Su3 U0 = link_at(a, b, p), U1 = link_at(a, up(a, b, p), q), U3 = link_at(a, up(a, b, q), p);
            Su3 w0 = mul_dag(U1, U3);
            Su3 w2 = mul_dag(w0, Ul);
            prod2alg(U0, w2, X);
            acc8(F, X, -r);
         }
         // (iv) base y - p: the link is U1
         b = dn(a, y, p);
         if (plaq_weight(a, n, a.x0[b], &r, &w1) && w1)
         {
            Su3 U0 = link_at(a, b, p), U2 = link_at(a, b, q), U3 = link_at(a, up(a, b, q), p);
            Su3 w0 = mul_dag(Ul, U3);
            Su3 wb = dag_mul(U2, U0);
            prod2alg(w0, wb, X);
            acc8(F, X, r);
         }
      }
   }
   double2 *f = a.frc + (size_t)(j >> 6) * (32 * SB) + (j & (SB - 1)) + (size_t)(4 * l) * SB;
#pragma unroll
   for (int c = 0; c < 4; c++) f[c * SB] = make_double2(F[2 * c], F[2 * c + 1]);
}

// ---- plaquette sums (uflds/plaq_sum.c:69-175) ----

// Re tr U(p) of the plaquette in plane n at point b: U(p) = U0 U1 U3^+ U2^+
// with the links of plaq_uidx (lattice/uidx.c:29-40), evaluated as
// Re tr{(U0 U1) (U2 U3)^+}
__device__ inline double plaq_re(const FArgs &a, int b, int n)
{
   int p, q;
   plane_pq(n, p, q);
   Su3 U0 = link_at(a, b, p), U1 = link_at(a, up(a, b, p), q);
   Su3 U2 = link_at(a, b, q), U3 = link_at(a, up(a, b, q), p);
   Su3 w1 = mul(U0, U1), w2 = mul(U2, U3);
   double s = 0.0;
#pragma unroll
   for (int k = 0; k < 9; k++) s += w1.e[k].x * w2.e[k].x + w1.e[k].y * w2.e[k].y;
   return s;
}

// per point: the time-like planes except at x0 = N0-1 with open bc, the
// space-like planes with weight wp on the boundary slices of open bc
// (plaq_sum.c:116-146 for boundary types 0 and 3); double-double sum per
// work-group
__global__ __launch_bounds__(256) void k_plaq_partial(FArgs a, double wp, dd *part)
{
   dd acc = {0.0, 0.0};
   for (int b = blockIdx.x * 256 + threadIdx.x; b < a.V; b += gridDim.x * 256)
   {
      const int t = a.x0[b];
      double pa = 0.0;
      if (t < a.N0 - 1 || a.bc != 0)
         for (int n = 0; n < 3; n++) pa += plaq_re(a, b, n);
      if ((t > 0 && t < a.N0 - 1) || a.bc == 3)
      {
         for (int n = 3; n < 6; n++) pa += plaq_re(a, b, n);
      }
      else
      {
         for (int n = 3; n < 6; n++) pa += wp * plaq_re(a, b, n);
      }
      acc = dd_add(acc, dd{pa, 0.0});
   }
   dd v[1] = {acc};
   block_reduce_dd<1>(v);
   if (threadIdx.x == 0) part[blockIdx.x] = v[0];
}

__global__ __launch_bounds__(256) void k_plaq_final(const dd *part, int nb, dd *out)
{
   dd v[1] = {{0.0, 0.0}};
   for (int k = threadIdx.x; k < nb; k += 256) v[0] = dd_add(v[0], part[k]);
   block_reduce_dd<1>(v);
   if (threadIdx.x == 0) out[0] = v[0];
}

// ---- SU(3) exponential, Cayley-Hamilton form (chexp.c) ----

struct ChexpCoef
{
   int N;
   double c[32];
};

namespace {

__device__ inline void alg_rows(const double *X, double2 v[3][3])
{
   v[0][0] = make_double2(0.0, X[0] + X[1]);
   v[0][1] = make_double2(X[2], X[3]);
   v[0][2] = make_double2(X[4], X[5]);
   v[1][0] = make_double2(-X[2], X[3]);
   v[1][1] = make_double2(0.0, X[1] - 2.0 * X[0]);
   v[1][2] = make_double2(X[6], X[7]);
   v[2][0] = make_double2(-X[4], X[5]);
   v[2][1] = make_double2(-X[6], X[7]);
   v[2][2] = make_double2(0.0, X[0] - 2.0 * X[1]);
}

__device__ inline double vre(const double2 *a, const double2 *b)
{
   return a[0].x * b[0].x + a[0].y * b[0].y + a[1].x * b[1].x + a[1].y * b[1].y + a[2].x * b[2].x + a[2].y * b[2].y;
}

__device__ inline double vim(const double2 *a, const double2 *b)
{
   return a[0].x * b[0].y - a[0].y * b[0].x + a[1].x * b[1].y - a[1].y * b[1].x + a[2].x * b[2].y - a[2].y * b[2].x;
}

// u = exp(X) for |X| small enough (ch2mat(chexp_drv0(X)))
__device__ inline Su3 expX(const double *X, const ChexpCoef &cc)
{
   double2 v[3][3], w[3];
   double t = 3.0 * (X[0] * X[0] + X[1] * X[1] - X[0] * X[1]) + X[2] * X[2] + X[3] * X[3] + X[4] * X[4] +
              X[5] * X[5] + X[6] * X[6] + X[7] * X[7];
   alg_rows(X, v);
   // w = v1 x v2 (_vector_cross_prod)
#pragma unroll
   for (int k = 0; k < 3; k++)
   {
      int i1 = (k + 1) % 3, i2 = (k + 2) % 3;
      double2 x1 = v[1][i1], x2 = v[1][i2], y1 = v[2][i2], y2 = v[2][i1];
      w[k] = make_double2(x1.x * y1.x - x1.y * y1.y - x2.x * y2.x + x2.y * y2.y,
                          x2.x * y2.y + x2.y * y2.x - x1.x * y1.y - x1.y * y1.x);
   }
   double d = vim(v[0], w);
   double p0r = cc.c[cc.N - 6], p0i = 0.0, p1r = 0.0, p1i = 0.0, p2r = 0.0, p2i = 0.0;
   for (int n = cc.N - 7; n >= 0; n--)
   {
      double q0r = p0r, q0i = p0i, q1r = p1r, q1i = p1i, q2r = p2r, q2i = p2i;
      p0r = cc.c[n] + d * q2i;
      p0i = -d * q2r;
      p1r = q0r - t * q2r;
      p1i = q0i - t * q2i;
      p2r = q1r;
      p2i = q1i;
   }
   Su3 u;
#pragma unroll
   for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++)
      {
         if (i == j)
            u.e[3 * i + j] = make_double2(p0r - p1i * v[i][j].y, p0i + p1r * v[i][j].y);
         else
            u.e[3 * i + j] = make_double2(p1r * v[i][j].x - p1i * v[i][j].y, p1r * v[i][j].y + p1i * v[i][j].x);
      }
#pragma unroll
   for (int i = 0; i < 3; i++)
   {
      double zr = vre(v[i], v[i]);
      u.e[4 * i].x -= p2r * zr;
      u.e[4 * i].y -= p2i * zr;
   }
   const int pr[3][2] = {{0, 1}, {0, 2}, {1, 2}};
#pragma unroll
   for (int m = 0; m < 3; m++)
   {
      int i = pr[m][0], j = pr[m][1];
      double zr = vre(v[i], v[j]), zi = vim(v[i], v[j]);
      u.e[3 * i + j].x -= p2r * zr + p2i * zi;
      u.e[3 * i + j].y -= p2i * zr - p2r * zi;
      u.e[3 * j + i].x -= p2r * zr - p2i * zi;
      u.e[3 * j + i].y -= p2i * zr + p2r * zi;
   }
   return u;
}

// u <- exp(eps X) u (expXsu3, chexp.c:878-929)
__device__ inline void exp_update(double eps, const double *X, Su3 &u, const ChexpCoef &cc)
{
   double nfrb = 4.0 * (3.0 * (X[0] * X[0] + X[1] * X[1] - X[0] * X[1]) + X[2] * X[2] + X[3] * X[3] +
                        X[4] * X[4] + X[5] * X[5] + X[6] * X[6] + X[7] * X[7]);
   nfrb *= eps * eps;
   int n = 0;
   while (nfrb > 3.0)
   {
      nfrb *= 0.25;
      eps *= 0.5;
      n += 1;
   }
   double Y[8];
#pragma unroll
   for (int k = 0; k < 8; k++) Y[k] = eps * X[k];
   Su3 e = expX(Y, cc);
   for (int k = 0; k < n; k++) e = mul(e, e);
   u = mul(e, u);
}

#pragma clang fp contract(on)

}  // namespace

// one lane per link of the odd points: update rules of wflow.c:96-144
__global__ __launch_bounds__(256) void k_update_ud(FArgs a, ChexpCoef cc)
{
   int k = blockIdx.x * 256 + threadIdx.x;
   if (k >= 4 * a.V) return;
   int j = k >> 3, l = k & 7, o = a.h + j;
   int t = a.x0[o];
   double e = a.eps;
   bool upd = true;
   if (t == 0)
   {
      if (l == 1)
         upd = (a.bc != 0);
      else if (l >= 2 && a.bc == 0)
         e = 2.0 * a.eps;
   }
   else if (t == a.N0 - 1)
   {
      if (l == 0)
         upd = (a.bc != 0);
      else if (l >= 2 && a.bc == 0)
         e = 2.0 * a.eps;
   }
   if (!upd) return;
   const double2 *f = a.frc + (size_t)(j >> 6) * (32 * SB) + (j & (SB - 1)) + (size_t)(4 * l) * SB;
   double X[8];
#pragma unroll
   for (int c = 0; c < 4; c++)
   {
      double2 v = f[c * SB];
      X[2 * c] = v.x;
      X[2 * c + 1] = v.y;
   }
   double2 *up = a.u_w + (size_t)(j >> 6) * (72 * SB) + (j & (SB - 1)) + (size_t)(9 * l) * SB;
   Su3 u;
#pragma unroll
   for (int c = 0; c < 9; c++) u.e[c] = up[c * SB];
   exp_update(e, X, u, cc);
#pragma unroll
   for (int c = 0; c < 9; c++) up[c * SB] = u.e[c];
}

// time-like links at x0 = N0-1 times -1 (periodic bc: antiperiodic quarks)
__global__ __launch_bounds__(256) void k_ud_phase(double2 *ud, const int *x0, int h, int N0)
{
   int k = blockIdx.x * 256 + threadIdx.x;
   if (k >= 2 * h) return;
   int j = k >> 1, l = k & 1, o = h + j, t = x0[o];
   // slot 0 is U(o,0) (at time t), slot 1 is U(o-0,0) (at time t-1)
   if (!((l == 0 && t == N0 - 1) || (l == 1 && t == 0))) return;
   double2 *p = ud + (size_t)(j >> 6) * (72 * SB) + (j & (SB - 1)) + (size_t)(9 * l) * SB;
#pragma unroll
   for (int c = 0; c < 9; c++) p[c * SB] = make_double2(-p[c * SB].x, -p[c * SB].y);
}

static ChexpCoef chexp_coef()
{
   ChexpCoef cc;
   int N = 7;
   double fact = 1.0;
   while (fact > DBL_EPSILON)
   {
      N += 1;
      fact /= (double)(N - 7);
   }
   N += (N % 2);
   cc.N = N;
   cc.c[0] = 1.0;
   for (int k = 0; k < N; k++) cc.c[k + 1] = cc.c[k] / (double)(k + 1);
   return cc;
}

static double2 *d_frc = nullptr;
static size_t frc_bytes = 0;

static FArgs fargs(bool need_frc = true)
{
   Lib &L = lib();
   FArgs a;
   a.ud = L.d_ud;
   a.u_w = L.d_ud;
   a.nbr = L.d_nbr;
   a.x0 = L.d_x0;
   a.V = L.g.V;
   a.h = L.g.V / 2;
   a.N0 = L.g.L[0];
   a.bc = L.bc;
   a.eps = 0.0;
   a.frc = nullptr;
   if (!need_frc) return a;
   size_t need = units_for(a.h, 32) * sizeof(double2);
   if (need > frc_bytes)
   {
      dev_free(d_frc);
      d_frc = (double2 *)dev_alloc(need);
      frc_bytes = d_frc ? need : 0;
   }
   a.frc = d_frc;
   return a;
}

static bool flow_ready(const char *where)
{
   Lib &L = lib();
   if (!require_geometry(where)) return false;
   if (!L.ud_set)
   {
      fail(where, "gauge field is not set");
      return false;
   }
   return true;
}

void release_flow()
{
   dev_free(d_frc);
   d_frc = nullptr;
   frc_bytes = 0;
}

}  // namespace oqh

using namespace oqh;

extern "C" void plaq_frc(void)
{
   Lib &L = lib();
   if (!flow_ready("plaq_frc [force0.c]")) return;
   FArgs a = fargs();
   if (!a.frc) return;
   unsigned nb = (unsigned)((4 * (size_t)a.V + 255) / 256);
   k_plaq_frc<<<nb, 256, 0, L.stream>>>(a);
   HIP_OK(hipStreamSynchronize(L.stream));
}

extern "C" void oqh_export_frc(double *frc)
{
   Lib &L = lib();
   if (!flow_ready("oqh_export_frc") || !d_frc) return;
   int h = L.g.V / 2;
   size_t bytes = (size_t)h * 8 * 8 * sizeof(double);
   void *aos = tmp_buffer(bytes);
   blocks_to_aos_f64(d_frc, (double2 *)aos, h, 32, L.stream);
   HIP_OK(hipMemcpyAsync(frc, aos, bytes, hipMemcpyDeviceToHost, L.stream));
   HIP_OK(hipStreamSynchronize(L.stream));
}

extern "C" void fwd_euler(int n, double eps)
{
   Lib &L = lib();
   if (!flow_ready("fwd_euler [wflow.c]") || n <= 0) return;
   FArgs a = fargs();
   if (!a.frc) return;
   ChexpCoef cc = chexp_coef();
   unsigned nb = (unsigned)((4 * (size_t)a.V + 255) / 256);
   a.eps = -eps;
   for (int k = 0; k < n; k++)
   {
      k_plaq_frc<<<nb, 256, 0, L.stream>>>(a);
      k_update_ud<<<nb, 256, 0, L.stream>>>(a, cc);
   }
   HIP_OK(hipStreamSynchronize(L.stream));
   L.ud_version++;
   L.u_valid = false;
   L.swd_valid = L.sw_valid = false;
}

extern "C" void set_ud_phase(void)
{
   Lib &L = lib();
   if (!flow_ready("set_ud_phase [uflds.c]") || L.ud_phase) return;
   if (L.bc == 3)
   {
      int h = L.g.V / 2;
      k_ud_phase<<<(unsigned)((2 * (size_t)h + 255) / 256), 256, 0, L.stream>>>(L.d_ud, L.d_x0, h, L.g.L[0]);
      HIP_OK(hipStreamSynchronize(L.stream));
      L.ud_version++;
      L.u_valid = false;
      L.swd_valid = L.sw_valid = false;
   }
   L.ud_phase = true;
}

extern "C" void unset_ud_phase(void)
{
   Lib &L = lib();
   if (!flow_ready("unset_ud_phase [uflds.c]") || !L.ud_phase) return;
   L.ud_phase = false;
   if (L.bc == 3)
   {
      int h = L.g.V / 2;
      k_ud_phase<<<(unsigned)((2 * (size_t)h + 255) / 256), 256, 0, L.stream>>>(L.d_ud, L.d_x0, h, L.g.L[0]);
      HIP_OK(hipStreamSynchronize(L.stream));
      L.ud_version++;
      L.u_valid = false;
      L.swd_valid = L.sw_valid = false;
   }
}

// uflds/plaq_sum.c:155-191: sums of Re tr U(p) over all unoriented plaquettes
// (plaq_wsum_dble: space-like plaquettes on the boundary slices of open bc
// with weight 1/2).  One process holds the whole lattice, so icom has no
// effect.
static double plaq_sum(double wp, const char *where)
{
   Lib &L = lib();
   if (!flow_ready(where)) return 0.0;
   static dd *d_part = nullptr, *h_out = nullptr;
   const int nbmax = 1024;
   if (!d_part) d_part = (dd *)dev_alloc(sizeof(dd) * (nbmax + 1));
   if (!h_out) HIP_OK(hipHostMalloc((void **)&h_out, sizeof(dd), 0));
   if (!d_part || !h_out) return 0.0;
   FArgs a = oqh::fargs(false);
   int nb = (a.V + 255) / 256;
   nb = nb > nbmax ? nbmax : nb;
   oqh::k_plaq_partial<<<nb, 256, 0, L.stream>>>(a, wp, d_part);
   oqh::k_plaq_final<<<1, 256, 0, L.stream>>>(d_part, nb, d_part + nbmax);
   HIP_OK(hipMemcpyAsync(h_out, d_part + nbmax, sizeof(dd), hipMemcpyDeviceToHost, L.stream));
   HIP_OK(hipStreamSynchronize(L.stream));
   return h_out->hi;
}

extern "C" double plaq_sum_dble(int icom)
{
   (void)icom;
   return plaq_sum(1.0, "plaq_sum_dble [plaq_sum.c]");
}

extern "C" double plaq_wsum_dble(int icom)
{
   (void)icom;
   return plaq_sum(0.5, "plaq_wsum_dble [plaq_sum.c]");
}
