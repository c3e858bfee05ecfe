// dirac_dev.h -- device building blocks of the Wilson-Dirac operator shared
// by the full-lattice kernels (dirac.hip) and the SAP block solver (sap.hip):
// site-block loads/stores, the spin projections (1 -+ gamma_mu) with the
// colour products of dirac/Dw_dble.c:1114-1373, and the Pauli matrix-vector
// product of sw_term/pauli_dble.c:1336-1440.
#pragma once

#include "oqh_internal.h"

namespace oqh {

template <typename C> __device__ inline void load_spinor(const C *f, int y, C *s)
{
   const C *p = f + (size_t)(y >> 6) * (12 * SB) + (y & (SB - 1));
#pragma unroll
   for (int c = 0; c < 12; c++) s[c] = p[c * SB];
}

template <typename C> __device__ inline void store_spinor(C *f, int y, const C *s)
{
   C *p = f + (size_t)(y >> 6) * (12 * SB) + (y & (SB - 1));
#pragma unroll
   for (int c = 0; c < 12; c++) p[c * SB] = s[c];
}

// load of data used once (NT: non-temporal, not kept in the caches)
template <bool NT, typename C> __device__ inline C ld_stream(const C *p)
{
   if (!NT) return *p;
   C v;
   v.x = __builtin_nontemporal_load(&p->x);
   v.y = __builtin_nontemporal_load(&p->y);
   return v;
}

// streaming store (non-temporal: the output is not read again by the kernel)
template <typename C> __device__ inline void store_spinor_nt(C *f, int y, const C *s)
{
   C *p = f + (size_t)(y >> 6) * (12 * SB) + (y & (SB - 1));
#pragma unroll
   for (int c = 0; c < 12; c++)
   {
      __builtin_nontemporal_store(s[c].x, &(p[c * SB].x));
      __builtin_nontemporal_store(s[c].y, &(p[c * SB].y));
   }
}

template <typename C> __device__ inline void zero_spinor(C *f, int y)
{
   C *p = f + (size_t)(y >> 6) * (12 * SB) + (y & (SB - 1));
   C z = cmk<C>(0, 0);
#pragma unroll
   for (int c = 0; c < 12; c++) p[c * SB] = z;
}

// c = U h  or  c = U^dag h  (U: 9 units starting at up, stride SB)
template <typename C, bool DAG> __device__ inline void su3_apply(const C *up, const C *h, C *c)
{
   C u[9];
#pragma unroll
   for (int k = 0; k < 9; k++) u[k] = up[k * SB];
#pragma unroll
   for (int i = 0; i < 3; i++)
   {
      if (!DAG)
         c[i] = cadd(cadd(cmul(u[3 * i], h[0]), cmul(u[3 * i + 1], h[1])), cmul(u[3 * i + 2], h[2]));
      else
         c[i] = cadd(cadd(cmulc(u[i], h[0]), cmulc(u[3 + i], h[1])), cmulc(u[6 + i], h[2]));
   }
}

// c = U h  or  c = U^dag h with U held in registers
template <typename C, bool DAG> __device__ inline void su3_mul(const C *u, const C *h, C *c)
{
#pragma unroll
   for (int i = 0; i < 3; i++)
   {
      if (!DAG)
         c[i] = cadd(cadd(cmul(u[3 * i], h[0]), cmul(u[3 * i + 1], h[1])), cmul(u[3 * i + 2], h[2]));
      else
         c[i] = cadd(cadd(cmulc(u[i], h[0]), cmulc(u[3 + i], h[1])), cmulc(u[6 + i], h[2]));
   }
}

// spin projection (1 -+ gamma_mu) s of the hopping term towards x+mu (FWD)
// or x-mu: the two independent colour vectors h0, h1
template <typename C, int MU, bool FWD> __device__ inline void spin_project(const C *sp, C *h0, C *h1)
{
   // spin components s_a are sp[3a..3a+2]
#pragma unroll
   for (int i = 0; i < 3; i++)
   {
      C s0 = sp[i], s1 = sp[3 + i], s2 = sp[6 + i], s3 = sp[9 + i];
      if (MU == 0) { h0[i] = FWD ? cadd(s0, s2) : csub(s0, s2); h1[i] = FWD ? cadd(s1, s3) : csub(s1, s3); }
      if (MU == 1) { h0[i] = FWD ? caddi(s0, s3) : csubi(s0, s3); h1[i] = FWD ? caddi(s1, s2) : csubi(s1, s2); }
      if (MU == 2) { h0[i] = FWD ? cadd(s0, s3) : csub(s0, s3); h1[i] = FWD ? csub(s1, s2) : cadd(s1, s2); }
      if (MU == 3) { h0[i] = FWD ? caddi(s0, s2) : csubi(s0, s2); h1[i] = FWD ? csubi(s1, s3) : caddi(s1, s3); }
   }
}

// acc += the full spinor rebuilt from the colour products c0, c1 of the
// projected components (inverse of spin_project)
template <typename C, int MU, bool FWD> __device__ inline void spin_expand_add(const C *c0, const C *c1, C *acc)
{
#pragma unroll
   for (int i = 0; i < 3; i++)
   {
      acc[i] = cadd(acc[i], c0[i]);
      acc[3 + i] = cadd(acc[3 + i], c1[i]);
      if (MU == 0)
      {
         acc[6 + i] = FWD ? cadd(acc[6 + i], c0[i]) : csub(acc[6 + i], c0[i]);
         acc[9 + i] = FWD ? cadd(acc[9 + i], c1[i]) : csub(acc[9 + i], c1[i]);
      }
      if (MU == 1)
      {
         acc[6 + i] = FWD ? csubi(acc[6 + i], c1[i]) : caddi(acc[6 + i], c1[i]);
         acc[9 + i] = FWD ? csubi(acc[9 + i], c0[i]) : caddi(acc[9 + i], c0[i]);
      }
      if (MU == 2)
      {
         acc[6 + i] = FWD ? csub(acc[6 + i], c1[i]) : cadd(acc[6 + i], c1[i]);
         acc[9 + i] = FWD ? cadd(acc[9 + i], c0[i]) : csub(acc[9 + i], c0[i]);
      }
      if (MU == 3)
      {
         acc[6 + i] = FWD ? csubi(acc[6 + i], c0[i]) : caddi(acc[6 + i], c0[i]);
         acc[9 + i] = FWD ? caddi(acc[9 + i], c1[i]) : csubi(acc[9 + i], c1[i]);
      }
   }
}

// one direction mu, forward (FWD: neighbour x+mu, link U(x,mu)) or backward
// (neighbour x-mu, link U(x-mu,mu)^dag): project s, multiply, expand
template <typename C, int MU, bool FWD>
__device__ inline void hop_one(const C *sp, const C *up, C *acc)
{
   C h0[3], h1[3], c0[3], c1[3];
   spin_project<C, MU, FWD>(sp, h0, h1);
   su3_apply<C, !FWD>(up, h0, c0);
   su3_apply<C, !FWD>(up, h1, c1);
   spin_expand_add<C, MU, FWD>(c0, c1, acc);
}

// link U at up into registers; REC: only rows 0 and 1 are read and row 2 is
// rebuilt as conj(row0 x row1), negated if neg (the links of an SU(3) field,
// some multiplied by -1 through the quark phase)
template <typename C, bool REC, bool NT = false>
__device__ inline void load_link(const C *up, bool neg, C *u)
{
   if (!REC)
   {
#pragma unroll
      for (int k = 0; k < 9; k++) u[k] = ld_stream<NT>(up + k * SB);
      return;
   }
#pragma unroll
   for (int k = 0; k < 6; k++) u[k] = ld_stream<NT>(up + k * SB);
   C c0 = csub(cmul(u[1], u[5]), cmul(u[2], u[4]));
   C c1 = csub(cmul(u[2], u[3]), cmul(u[0], u[5]));
   C c2 = csub(cmul(u[0], u[4]), cmul(u[1], u[3]));
   typedef decltype(C::x) R;
   const R sg = neg ? (R)-1 : (R)1;
   u[6] = cmk<C>(sg * c0.x, -sg * c0.y);
   u[7] = cmk<C>(sg * c1.x, -sg * c1.y);
   u[8] = cmk<C>(sg * c2.x, -sg * c2.y);
}

// address of link l (0..7, udfld order) stored with odd point j = y - V/2
template <typename C> __device__ inline const C *link_ptr(const C *u, int j, int l)
{
   return u + (size_t)(j >> 6) * (72 * SB) + (j & (SB - 1)) + (size_t)(9 * l) * SB;
}

// out = (A + i mu) v on one Weyl half (6 complex); m points at unit 0 of the
// half (stride SB); units 0-2 hold the real diagonal, 3-17 the upper
// off-diagonal entries (0,1),(0,2),...,(4,5)
template <typename C, bool NT = false>
__device__ inline void pauli_half(const C *m, decltype(C::x) mu, const C *v, C *out)
{
   C d0 = ld_stream<NT>(m), d1 = ld_stream<NT>(m + SB), d2 = ld_stream<NT>(m + 2 * SB);
   decltype(C::x) d[6] = {d0.x, d0.y, d1.x, d1.y, d2.x, d2.y};
#pragma unroll
   for (int i = 0; i < 6; i++) out[i] = cmk<C>(d[i] * v[i].x - mu * v[i].y, d[i] * v[i].y + mu * v[i].x);
   int p = 3;
#pragma unroll
   for (int i = 0; i < 6; i++)
#pragma unroll
      for (int j = i + 1; j < 6; j++)
      {
         C hij = ld_stream<NT>(m + p * SB);
         out[i] = cadd(out[i], cmul(hij, v[j]));
         out[j] = cadd(out[j], cmulc(hij, v[i]));
         p++;
      }
}

template <typename C, bool NT = false>
__device__ inline void pauli_site(const C *sw, int x, decltype(C::x) mu, const C *v, C *out)
{
   const C *m = sw + (size_t)(x >> 6) * (36 * SB) + (x & (SB - 1));
   pauli_half<C, NT>(m, mu, v, out);
   pauli_half<C, NT>(m + 18 * SB, -mu, v + 6, out + 6);
}

// out = (A + i mu gamma5-sign) v with A of one point given in the S-form:
// per Weyl half A = [[S3 + c, S1 - i S2], [S1 + i S2, -S3 + c]] with
// Hermitian 3x3 S_k (units 0-8: real diagonals of S1..S3 of both halves,
// units 9 + 9 half + 3 k + q: entries (0,1), (0,2), (1,2) of S_{k+1}); c is
// the diagonal constant of the point; the upper half uses +mu, the lower -mu
template <typename C>
__device__ inline void herm3_mul(const decltype(C::x) *d, const C *o, const C *v, C *r)
{
   // r = S v, S = [[d0, o0, o1], [o0*, d1, o2], [o1*, o2*, d2]]
   r[0] = cadd(cadd(cscale(v[0], d[0]), cmul(o[0], v[1])), cmul(o[1], v[2]));
   r[1] = cadd(cadd(cmulc(o[0], v[0]), cscale(v[1], d[1])), cmul(o[2], v[2]));
   r[2] = cadd(cadd(cmulc(o[1], v[0]), cmulc(o[2], v[1])), cscale(v[2], d[2]));
}

template <typename C, bool NT = false>
__device__ inline void sws_site(const C *sws, int x, decltype(C::x) c, decltype(C::x) mu, const C *v, C *out)
{
   typedef decltype(C::x) R;
   const C *m = sws + (size_t)(x >> 6) * (27 * SB) + (x & (SB - 1));
#pragma unroll
   for (int h = 0; h < 2; h++)
   {
      const R mm = h ? -mu : mu;
      const C *vu = v + 6 * h, *vd = v + 6 * h + 3;
      C *ou = out + 6 * h, *odn = out + 6 * h + 3;
#pragma unroll
      for (int i = 0; i < 3; i++)
      {
         ou[i] = cmk<C>(c * vu[i].x - mm * vu[i].y, c * vu[i].y + mm * vu[i].x);
         odn[i] = cmk<C>(c * vd[i].x - mm * vd[i].y, c * vd[i].y + mm * vd[i].x);
      }
      // S1, S2, S3 of this half one after the other, each used twice
#pragma unroll
      for (int k = 0; k < 3; k++)
      {
         R d[3];
#pragma unroll
         for (int i = 0; i < 3; i++)
         {
            const int g = 9 * h + 3 * k + i;  // among the 18 diagonal reals
            C w = ld_stream<NT>(m + (g >> 1) * SB);
            d[i] = (g & 1) ? w.y : w.x;
         }
         C o[3];
#pragma unroll
         for (int q = 0; q < 3; q++) o[q] = ld_stream<NT>(m + (9 + 9 * h + 3 * k + q) * SB);
         C a[3], b[3];
         herm3_mul(d, o, vd, a);  // S_k vd
         herm3_mul(d, o, vu, b);  // S_k vu
#pragma unroll
         for (int i = 0; i < 3; i++)
         {
            if (k == 0)
            {
               ou[i] = cadd(ou[i], a[i]);   // S1 vd
               odn[i] = cadd(odn[i], b[i]);  // S1 vu
            }
            else if (k == 1)
            {
               ou[i] = csubi(ou[i], a[i]);   // -i S2 vd
               odn[i] = caddi(odn[i], b[i]);  // +i S2 vu
            }
            else
            {
               ou[i] = cadd(ou[i], b[i]);   // S3 vu
               odn[i] = csub(odn[i], a[i]);  // -S3 vd
            }
         }
      }
   }
}

}  // namespace oqh
