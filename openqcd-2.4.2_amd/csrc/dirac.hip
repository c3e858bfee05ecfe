// dirac.hip -- O(a)-improved Wilson-Dirac operator on gfx950.
//
// Replaces dirac/Dw_dble.c:1377-1888 and dirac/Dw.c:973-1474.  The reference
// visits the odd points, gathers the hopping terms from the even neighbours
// (doe) and scatters the conjugate terms to them (deo), with a barrier
// schedule over 16 sub-blocks to keep the scatter race free.  Here every
// output point gathers its eight hopping terms itself, so no point is written
// by two work-items and no atomics or barriers are needed:
//
//   (D s)(x) = A(x) s(x) - 1/2 sum_mu [ (1-gamma_mu) U(x,mu) s(x+mu)
//                                      + (1+gamma_mu) U(x-mu,mu)^dag s(x-mu) ]
//
// Each (1 -+ gamma_mu) projection is applied before the colour product, so
// that only two colour vectors are multiplied per direction (the spin
// structure is that of the reference's doe/deo, Dw_dble.c:1114-1373).  A
// work-group of 256 handles the 128 even and the 128 odd points with the same
// offsets in the two halves of the site order; these lie in the same cache
// block, so the links and spinors shared by them are re-read from L2.
//
// Boundary rules (open bc): odd points on the boundary time slices are zeroed
// in input and output of Dw and contribute no hopping terms to even points
// (Dw_dble.c:1425-1456); the neighbour table carries that flag in bit 31.
#include <cfloat>
#include <cstdio>
#include <cstring>

#include <algorithm>

#include "dirac_dev.h"

namespace oqh {

enum DiracOp { OP_DW = 0, OP_DOE, OP_DEO, OP_HAT1, OP_HAT2, OP_DIAG };

template <typename T> struct DArgs
{
   typedef typename cplx_of<T>::t C;
   const C *s;     // input field
   C *r;           // output field
   C *s_w;         // input field, writable (zeroing side effects) or null
   const C *u;     // links, 72 units per odd point
   const C *sw;    // Pauli matrices, 36 units per point
   C *t;           // odd intermediate of the Schur complement
   const int *nbr; // [8][V]
   const unsigned char *bnd;
   int V, h;
   int p0, np;     // parity ops: first point and number of points
   T mu, muo, coef;
   int masked;     // 1 if boundary points exist (bc != 3)
   const unsigned char *lsign;  // per odd point, bit l: link l is -SU(3) (rebuilt links)
   const C *sws;   // SW term in the S-form (27 units per point) or null
   int sws_par;    // parities whose S-form term is usable (bit 0 even, bit 1 odd)
   const int *x0;  // time coordinates (diagonal constant of the S-form)
   int N0, bc;
   T cx1, cx3a, cx3b;  // 4 + m0 in the bulk, + cF - 1 next to open boundaries
   int xcdc;           // tiles per XCD run (xcd_tile)
};


template <typename T, int MU>
__device__ inline void hop_mu(const DArgs<T> &a, const typename cplx_of<T>::t *in, int x, int odd,
                              bool mask, typename cplx_of<T>::t *acc)
{
   typedef typename cplx_of<T>::t C;
   C sp[12];
   // x + mu
   int n = a.nbr[(size_t)(2 * MU) * a.V + x];
   if (!(mask && n < 0))
   {
      int y = n & 0x7fffffff;
      load_spinor(in, y, sp);
      const C *up = odd ? link_ptr(a.u, x - a.h, 2 * MU) : link_ptr(a.u, y - a.h, 2 * MU + 1);
      hop_one<C, MU, true>(sp, up, acc);
   }
   // x - mu
   n = a.nbr[(size_t)(2 * MU + 1) * a.V + x];
   if (!(mask && n < 0))
   {
      int y = n & 0x7fffffff;
      load_spinor(in, y, sp);
      const C *up = odd ? link_ptr(a.u, x - a.h, 2 * MU + 1) : link_ptr(a.u, y - a.h, 2 * MU);
      hop_one<C, MU, false>(sp, up, acc);
   }
}

template <typename T>
__device__ inline void hop_all(const DArgs<T> &a, const typename cplx_of<T>::t *in, int x, int odd,
                               bool mask, typename cplx_of<T>::t *acc)
{
   hop_mu<T, 0>(a, in, x, odd, mask, acc);
   hop_mu<T, 1>(a, in, x, odd, mask, acc);
   hop_mu<T, 2>(a, in, x, odd, mask, acc);
   hop_mu<T, 3>(a, in, x, odd, mask, acc);
}

// Work-groups are dispatched round-robin over the 8 XCDs of the chip, each
// with its own L2.  The tile of work-group b: with C > 0 each XCD takes runs
// of C consecutive tiles (C = 1 is the plain dispatch order), with C = 0 one
// contiguous range of tiles.  Consecutive tiles are neighbours within a
// thread block of the site order, so a run shares its face lines in one L2.
__device__ inline int xcd_tile(int b, int nb, int C)
{
   if (C == 0)
   {
      int q = nb >> 3, r = nb & 7, j = b & 7;
      return j * q + min(j, r) + (b >> 3);
   }
   const int span = 8 * C;
   if (C <= 1 || b >= nb / span * span) return b;
   const int g = b / span, rr = b % span;
   return g * span + (rr & 7) * C + (rr >> 3);
}

#ifndef OQH_DW_WAVES
#define OQH_DW_WAVES 0
#endif
#ifndef OQH_NT_STORE
#define OQH_NT_STORE 0
#endif
#if OQH_DW_WAVES
#define DW_ATTR __attribute__((amdgpu_waves_per_eu(OQH_DW_WAVES)))
#else
#define DW_ATTR
#endif

template <typename T, int OP>
__global__ __launch_bounds__(256) DW_ATTR void k_dirac(DArgs<T> a)
{
   typedef typename cplx_of<T>::t C;
   int x, odd;
   int b = xcd_tile(blockIdx.x, gridDim.x, a.xcdc);
   if (OP == OP_DW)
   {
      int tid = threadIdx.x;
      int k = b * 128 + (tid & 127);
      if (k >= a.h) return;
      odd = tid >> 7;
      x = k + odd * a.h;
   }
   else
   {
      int k = b * 256 + threadIdx.x;
      if (k >= a.np) return;
      x = a.p0 + k;
      odd = x >= a.h;
   }

   C acc[12];
#pragma unroll
   for (int c = 0; c < 12; c++) acc[c] = cmk<C>(0, 0);

   if (OP == OP_DW)
   {
      if (a.masked && a.bnd[x])
      {
         zero_spinor(a.r, x);
         zero_spinor(a.s_w, x);
         return;
      }
      hop_all(a, a.s, x, odd, a.masked != 0, acc);
      C v[12], o[12];
      load_spinor(a.s, x, v);
      pauli_site(a.sw, x, odd ? a.muo : a.mu, v, o);
#pragma unroll
      for (int c = 0; c < 12; c++) o[c] = cadd(o[c], cscale(acc[c], a.coef));
#if OQH_NT_STORE
      store_spinor_nt(a.r, x, o);
#else
      store_spinor(a.r, x, o);
#endif
   }
   else if (OP == OP_DOE)
   {
      if (a.masked && a.bnd[x])
      {
         zero_spinor(a.r, x);
         return;
      }
      hop_all(a, a.s, x, 1, a.masked != 0, acc);
#pragma unroll
      for (int c = 0; c < 12; c++) acc[c] = cscale(acc[c], a.coef);
      store_spinor(a.r, x, acc);
   }
   else if (OP == OP_DEO)
   {
      if (a.masked && a.bnd[x])
      {
         zero_spinor(a.r, x);
         return;
      }
      hop_all(a, a.s, x, 0, a.masked != 0, acc);
      C o[12];
      load_spinor((const C *)a.r, x, o);
#pragma unroll
      for (int c = 0; c < 12; c++) o[c] = cadd(o[c], cscale(acc[c], a.coef));
      store_spinor(a.r, x, o);
   }
   else if (OP == OP_HAT1)
   {
      if (a.masked && a.bnd[x])
      {
         zero_spinor(a.t, x);
         return;
      }
      hop_all(a, a.s, x, 1, a.masked != 0, acc);
      C o[12];
#pragma unroll
      for (int c = 0; c < 12; c++) acc[c] = cscale(acc[c], a.coef);
      pauli_site(a.sw, x, (T)0, acc, o);
      store_spinor(a.t, x, o);
   }
   else if (OP == OP_HAT2)
   {
      if (a.masked && a.bnd[x])
      {
         zero_spinor(a.r, x);
         return;
      }
      hop_all(a, (const C *)a.t, x, 0, a.masked != 0, acc);
      C v[12], o[12];
      load_spinor(a.s, x, v);
      pauli_site(a.sw, x, a.mu, v, o);
#pragma unroll
      for (int c = 0; c < 12; c++) o[c] = cadd(o[c], cscale(acc[c], a.coef));
      store_spinor(a.r, x, o);
   }
   else  // OP_DIAG
   {
      if (a.masked && a.bnd[x])
      {
         zero_spinor(a.r, x);
         if (a.s_w) zero_spinor(a.s_w, x);
         return;
      }
      C v[12], o[12];
      load_spinor(a.s, x, v);
      pauli_site(a.sw, x, a.mu, v, o);
      store_spinor(a.r, x, o);
   }
}

// ---------------------------------------------------------------------------
// Dw / Dw_dble with link reuse in registers and spinor reuse in LDS.
//
// A work-group of TP lanes takes TP even points e0..e0+TP-1 and the TP odd
// points with the same offsets; lane t owns the pair (e0+t, V/2+e0+t).  In
// the site order these tiles are compact (a 4^4 sub-block for 8^4 thread
// blocks), so 3/4 of all neighbours lie in the tile.  Every link is stored
// with an odd point and connects it to one even point: the owning lane loads
// it once and computes both hopping terms that use it, its own and the even
// neighbour's.  The even neighbour's term (two colour vectors after the spin
// projection) is passed through LDS; in each of the eight direction phases
// every even point receives at most one such term, so no two lanes write the
// same slot.  Even spinors are staged in LDS for the odd lanes.  Neighbours
// outside the tile are gathered from global memory as in k_dirac.  Compared
// with the plain gather this halves the link reads and removes 3/4 of the
// neighbour-spinor reads from the memory system.
// ---------------------------------------------------------------------------

#ifndef OQH_DW_TPD
#define OQH_DW_TPD 256
#endif
#ifndef OQH_DW_TPF
#define OQH_DW_TPF 256
#endif
#ifndef OQH_NT_LINK
#define OQH_NT_LINK 0
#endif
#ifndef OQH_NT_OUT
#define OQH_NT_OUT 1
#endif
#ifndef OQH_NT_SW
#define OQH_NT_SW 1
#endif
template <typename T> struct TileSize { static constexpr int n = OQH_DW_TPD; };
template <> struct TileSize<float> { static constexpr int n = OQH_DW_TPF; };
#ifndef OQH_DW_TILE
#define OQH_DW_TILE 1
#endif
#ifndef OQH_HAT_TILE
#define OQH_HAT_TILE 1
#endif

#ifndef OQH_NBR_PRELOAD
#define OQH_NBR_PRELOAD 1
#endif
#ifndef OQH_X_DOUBLE
#define OQH_X_DOUBLE 0
#endif
constexpr int XBUF = OQH_X_DOUBLE ? 2 : 1;

// neighbour-table entry k (0..7 for the odd point, 8..15 for the even point)
template <typename T> struct TileNbr
{
   int v[OQH_NBR_PRELOAD ? 16 : 1];
};

template <typename T, int K>
__device__ inline int tile_nbr(const DArgs<T> &a, const TileNbr<T> &nb, int e, int o)
{
#if OQH_NBR_PRELOAD
   (void)a; (void)e; (void)o;
   return nb.v[K];
#else
   (void)nb;
   return K < 8 ? a.nbr[(size_t)K * a.V + o] : a.nbr[(size_t)(K - 8) * a.V + e];
#endif
}

// output spinor of a tile kernel: non-temporal stores for the FP32
// operators (OQH_NT_OUT; Dw 4844 -> 4994 GB/s and Dwhat 3444 -> 3604 GB/s at
// 32^3x64), ordinary stores for FP64, whose kernels lost 0.5-0.9 % with them
template <typename C> __device__ inline void store_out(C *f, int y, const C *s)
{
   if (OQH_NT_OUT && sizeof(C) == sizeof(float2))
      store_spinor_nt(f, y, s);
   else
      store_spinor(f, y, s);
}

// OWN: the odd lane adds its own hopping term (from the even neighbour, S or
// global a.s) to acc_o; XCH: the odd lane passes the term of its even
// neighbour through X, and even lanes gather the terms of neighbours outside
// the tile from the odd-point field zsrc
template <typename T, int MU, bool FWD, int TP, bool OWN = true, bool XCH = true, bool REC = false>
__device__ inline void tile_phase(const DArgs<T> &a, int t, int e0, bool valid, const TileNbr<T> &nb,
                                  typename cplx_of<T>::t (*S)[TP], typename cplx_of<T>::t (*X)[TP],
                                  const typename cplx_of<T>::t *so, typename cplx_of<T>::t *acc_o,
                                  typename cplx_of<T>::t *acc_e, const typename cplx_of<T>::t *zsrc)
{
   typedef typename cplx_of<T>::t C;
   const int e = e0 + t, o = a.h + e;
   // odd point o: link of direction (MU, FWD), neighbour y = o +- mu (even)
   if (valid)
   {
      int ny = tile_nbr<T, 2 * MU + (FWD ? 0 : 1)>(a, nb, e, o);
      int y = ny & 0x7fffffff;
      int ly = y - e0;
      bool yin = (unsigned)ly < (unsigned)TP;
      const C *up = link_ptr(a.u, e, FWD ? 2 * MU : 2 * MU + 1);
      C u[9];
      load_link<C, REC, OQH_NT_LINK>(up, REC && ((a.lsign[e] >> (FWD ? 2 * MU : 2 * MU + 1)) & 1), u);
      C h0[3], h1[3], c0[3], c1[3];
      if (OWN && !(a.masked && ny < 0))
      {
         C sy[12];
         if (yin)
         {
#pragma unroll
            for (int c = 0; c < 12; c++) sy[c] = S[c][ly];
         }
         else
            load_spinor(a.s, y, sy);
         spin_project<C, MU, FWD>(sy, h0, h1);
         su3_mul<C, !FWD>(u, h0, c0);
         su3_mul<C, !FWD>(u, h1, c1);
         spin_expand_add<C, MU, FWD>(c0, c1, acc_o);
      }
      if (XCH && yin)
      {
         // the term of y from o, as y's neighbour in the opposite direction
         spin_project<C, MU, !FWD>(so, h0, h1);
         su3_mul<C, FWD>(u, h0, c0);
         su3_mul<C, FWD>(u, h1, c1);
#pragma unroll
         for (int c = 0; c < 3; c++)
         {
            X[c][ly] = c0[c];
            X[3 + c][ly] = c1[c];
         }
      }
   }
   if (!XCH) return;
   __syncthreads();
   // even point e: neighbour z = e -+ mu (odd), the opposite direction
   if (valid)
   {
      int nz = tile_nbr<T, 8 + 2 * MU + (FWD ? 1 : 0)>(a, nb, e, o);
      if (!(a.masked && nz < 0))
      {
         int z = nz & 0x7fffffff;
         if ((unsigned)(z - a.h - e0) < (unsigned)TP)
         {
            C c0[3], c1[3];
#pragma unroll
            for (int c = 0; c < 3; c++)
            {
               c0[c] = X[c][t];
               c1[c] = X[3 + c][t];
            }
            spin_expand_add<C, MU, !FWD>(c0, c1, acc_e);
         }
         else
         {
            C sz[12], u[9], h0[3], h1[3], c0[3], c1[3];
            load_spinor(zsrc, z, sz);
            const int l = FWD ? 2 * MU : 2 * MU + 1;
            load_link<C, REC>(link_ptr(a.u, z - a.h, l), REC && ((a.lsign[z - a.h] >> l) & 1), u);
            spin_project<C, MU, !FWD>(sz, h0, h1);
            su3_mul<C, FWD>(u, h0, c0);
            su3_mul<C, FWD>(u, h1, c1);
            spin_expand_add<C, MU, !FWD>(c0, c1, acc_e);
         }
      }
   }
   // with a single exchange buffer the next phase may overwrite it only
   // after every lane has read it; with two buffers the next phase's barrier
   // gives that order
   if (XBUF == 1) __syncthreads();
}

// diagonal constant of the SW term at point x (sw_term.c:293-309)
template <typename T> __device__ inline T sw_const(const DArgs<T> &a, int x)
{
   if (a.bc == 3) return a.cx1;
   const int t = a.x0[x];
   if (t == 0 || (t == a.N0 - 1 && a.bc == 0)) return (T)1;
   if (t == 1) return a.cx3a;
   if (t == a.N0 - 2 && a.bc == 0) return a.cx3b;
   return a.cx1;
}

// (A + i mu) v at point x from the Pauli field or from the S-form
template <typename T, bool SWS>
__device__ inline void site_term(const DArgs<T> &a, int x, T mu, const typename cplx_of<T>::t *v,
                                 typename cplx_of<T>::t *out)
{
   if (SWS)
      sws_site<typename cplx_of<T>::t, OQH_NT_SW>(a.sws, x, sw_const(a, x), mu, v, out);
   else
      pauli_site<typename cplx_of<T>::t, OQH_NT_SW>(a.sw, x, mu, v, out);
}

template <typename T, int TP = TileSize<T>::n, bool REC = false, bool SWS = false>
__global__ __launch_bounds__(TP) void k_dw_tile(DArgs<T> a)
{
   typedef typename cplx_of<T>::t C;
   __shared__ C S[12][TP];
   __shared__ C X[XBUF][6][TP];
   const int t = threadIdx.x;
   const int e0 = xcd_tile(blockIdx.x, gridDim.x, a.xcdc) * TP;
   const int e = e0 + t, o = a.h + e;
   const bool valid = e < a.h;
   bool eb = false, ob = false;
   TileNbr<T> nb;
   C so[12], acc_o[12], acc_e[12];
#pragma unroll
   for (int c = 0; c < 12; c++)
   {
      acc_o[c] = cmk<C>(0, 0);
      acc_e[c] = cmk<C>(0, 0);
      so[c] = cmk<C>(0, 0);
   }
   if (valid)
   {
#if OQH_NBR_PRELOAD
#pragma unroll
      for (int k = 0; k < 8; k++)
      {
         nb.v[k] = a.nbr[(size_t)k * a.V + o];
         nb.v[8 + k] = a.nbr[(size_t)k * a.V + e];
      }
#endif
      eb = a.masked && a.bnd[e];
      ob = a.masked && a.bnd[o];
      C se[12];
      if (eb)
      {
#pragma unroll
         for (int c = 0; c < 12; c++) se[c] = cmk<C>(0, 0);
      }
      else
         load_spinor(a.s, e, se);
#pragma unroll
      for (int c = 0; c < 12; c++) S[c][t] = se[c];
      if (!ob) load_spinor(a.s, o, so);
   }
   __syncthreads();
   tile_phase<T, 0, true, TP, true, true, REC>(a, t, e0, valid, nb, S, X[0], so, acc_o, acc_e, a.s);
   tile_phase<T, 0, false, TP, true, true, REC>(a, t, e0, valid, nb, S, X[XBUF - 1], so, acc_o, acc_e, a.s);
   tile_phase<T, 1, true, TP, true, true, REC>(a, t, e0, valid, nb, S, X[0], so, acc_o, acc_e, a.s);
   tile_phase<T, 1, false, TP, true, true, REC>(a, t, e0, valid, nb, S, X[XBUF - 1], so, acc_o, acc_e, a.s);
   tile_phase<T, 2, true, TP, true, true, REC>(a, t, e0, valid, nb, S, X[0], so, acc_o, acc_e, a.s);
   tile_phase<T, 2, false, TP, true, true, REC>(a, t, e0, valid, nb, S, X[XBUF - 1], so, acc_o, acc_e, a.s);
   tile_phase<T, 3, true, TP, true, true, REC>(a, t, e0, valid, nb, S, X[0], so, acc_o, acc_e, a.s);
   tile_phase<T, 3, false, TP, true, true, REC>(a, t, e0, valid, nb, S, X[XBUF - 1], so, acc_o, acc_e, a.s);
   if (!valid) return;
   C out[12];
   // boundary points of open bc are zeroed in output and input
   if (ob)
   {
      zero_spinor(a.r, o);
      zero_spinor(a.s_w, o);
   }
   else
   {
      site_term<T, SWS>(a, o, a.muo, so, out);
#pragma unroll
      for (int c = 0; c < 12; c++) out[c] = cadd(out[c], cscale(acc_o[c], a.coef));
      store_out(a.r, o, out);
   }
   if (eb)
   {
      zero_spinor(a.r, e);
      zero_spinor(a.s_w, e);
   }
   else
   {
      C se[12];
#pragma unroll
      for (int c = 0; c < 12; c++) se[c] = S[c][t];
      site_term<T, SWS>(a, e, a.mu, se, out);
#pragma unroll
      for (int c = 0; c < 12; c++) out[c] = cadd(out[c], cscale(acc_e[c], a.coef));
      store_out(a.r, e, out);
   }
}

// Dwhat / Dwhat_dble in two tile passes (Dw_dble.c:1792-1888):
//   pass 1 (odd points):  t_o = A_oo^-1 (-1/2 sum of hops from s_e), the even
//                         spinors of the tile staged in LDS;
//   pass 2 (even points): r_e = (A_ee + i mu g5) s_e + 1/2 sum of hops from
//                         t_o, each link loaded once by its odd lane, which
//                         passes the even neighbour's term through LDS.
template <typename T, int TP = TileSize<T>::n, bool REC = false>
__global__ __launch_bounds__(TP) void k_hat1_tile(DArgs<T> a)
{
   typedef typename cplx_of<T>::t C;
   __shared__ C S[12][TP];
   const int t = threadIdx.x;
   const int e0 = xcd_tile(blockIdx.x, gridDim.x, a.xcdc) * TP;
   const int e = e0 + t, o = a.h + e;
   const bool valid = e < a.h;
   bool ob = false;
   TileNbr<T> nb;
   C acc[12];
#pragma unroll
   for (int c = 0; c < 12; c++) acc[c] = cmk<C>(0, 0);
   if (valid)
   {
#if OQH_NBR_PRELOAD
#pragma unroll
      for (int k = 0; k < 8; k++) nb.v[k] = a.nbr[(size_t)k * a.V + o];
#endif
      ob = a.masked && a.bnd[o];
      C se[12];
      if (a.masked && a.bnd[e])
      {
#pragma unroll
         for (int c = 0; c < 12; c++) se[c] = cmk<C>(0, 0);
      }
      else
         load_spinor(a.s, e, se);
#pragma unroll
      for (int c = 0; c < 12; c++) S[c][t] = se[c];
   }
   __syncthreads();
   tile_phase<T, 0, true, TP, true, false, REC>(a, t, e0, valid, nb, S, nullptr, nullptr, acc, nullptr, nullptr);
   tile_phase<T, 0, false, TP, true, false, REC>(a, t, e0, valid, nb, S, nullptr, nullptr, acc, nullptr, nullptr);
   tile_phase<T, 1, true, TP, true, false, REC>(a, t, e0, valid, nb, S, nullptr, nullptr, acc, nullptr, nullptr);
   tile_phase<T, 1, false, TP, true, false, REC>(a, t, e0, valid, nb, S, nullptr, nullptr, acc, nullptr, nullptr);
   tile_phase<T, 2, true, TP, true, false, REC>(a, t, e0, valid, nb, S, nullptr, nullptr, acc, nullptr, nullptr);
   tile_phase<T, 2, false, TP, true, false, REC>(a, t, e0, valid, nb, S, nullptr, nullptr, acc, nullptr, nullptr);
   tile_phase<T, 3, true, TP, true, false, REC>(a, t, e0, valid, nb, S, nullptr, nullptr, acc, nullptr, nullptr);
   tile_phase<T, 3, false, TP, true, false, REC>(a, t, e0, valid, nb, S, nullptr, nullptr, acc, nullptr, nullptr);
   if (!valid) return;
   if (ob)
   {
      zero_spinor(a.t, o);
      return;
   }
   C out[12];
#pragma unroll
   for (int c = 0; c < 12; c++) acc[c] = cscale(acc[c], a.coef);
   pauli_site<C, OQH_NT_SW>(a.sw, o, (T)0, acc, out);
   store_out(a.t, o, out);
}

template <typename T, int TP = TileSize<T>::n, bool REC = false, bool SWS = false>
__global__ __launch_bounds__(TP) void k_hat2_tile(DArgs<T> a)
{
   typedef typename cplx_of<T>::t C;
   __shared__ C X[XBUF][6][TP];
   const int t = threadIdx.x;
   const int e0 = xcd_tile(blockIdx.x, gridDim.x, a.xcdc) * TP;
   const int e = e0 + t, o = a.h + e;
   const bool valid = e < a.h;
   bool eb = false;
   TileNbr<T> nb;
   C to[12], acc[12];
#pragma unroll
   for (int c = 0; c < 12; c++)
   {
      acc[c] = cmk<C>(0, 0);
      to[c] = cmk<C>(0, 0);
   }
   if (valid)
   {
#if OQH_NBR_PRELOAD
#pragma unroll
      for (int k = 0; k < 8; k++)
      {
         nb.v[k] = a.nbr[(size_t)k * a.V + o];
         nb.v[8 + k] = a.nbr[(size_t)k * a.V + e];
      }
#endif
      eb = a.masked && a.bnd[e];
      if (!(a.masked && a.bnd[o])) load_spinor((const C *)a.t, o, to);
   }
   const C *tz = a.t;
   tile_phase<T, 0, true, TP, false, true, REC>(a, t, e0, valid, nb, nullptr, X[0], to, nullptr, acc, tz);
   tile_phase<T, 0, false, TP, false, true, REC>(a, t, e0, valid, nb, nullptr, X[XBUF - 1], to, nullptr, acc, tz);
   tile_phase<T, 1, true, TP, false, true, REC>(a, t, e0, valid, nb, nullptr, X[0], to, nullptr, acc, tz);
   tile_phase<T, 1, false, TP, false, true, REC>(a, t, e0, valid, nb, nullptr, X[XBUF - 1], to, nullptr, acc, tz);
   tile_phase<T, 2, true, TP, false, true, REC>(a, t, e0, valid, nb, nullptr, X[0], to, nullptr, acc, tz);
   tile_phase<T, 2, false, TP, false, true, REC>(a, t, e0, valid, nb, nullptr, X[XBUF - 1], to, nullptr, acc, tz);
   tile_phase<T, 3, true, TP, false, true, REC>(a, t, e0, valid, nb, nullptr, X[0], to, nullptr, acc, tz);
   tile_phase<T, 3, false, TP, false, true, REC>(a, t, e0, valid, nb, nullptr, X[XBUF - 1], to, nullptr, acc, tz);
   if (!valid) return;
   if (eb)
   {
      zero_spinor(a.r, e);
      return;
   }
   C se[12], out[12];
   load_spinor(a.s, e, se);
   site_term<T, SWS>(a, e, a.mu, se, out);
#pragma unroll
   for (int c = 0; c < 12; c++) out[c] = cadd(out[c], cscale(acc[c], a.coef));
   store_out(a.r, e, out);
}

// ---------------------------------------------------------------------------
// Dw / Dw_dble on several fields at once (no reference counterpart): the
// tile kernel with NR lanes per even-odd pair, one per right-hand side.  The
// NR lanes of a pair are adjacent in a wavefront and load the same links and
// Pauli matrices, so each such load is fetched once for all right-hand sides;
// the spinor staging and exchange are per right-hand side.
// ---------------------------------------------------------------------------

constexpr int NRMAX = 8;

template <typename T> struct MArgs
{
   typedef typename cplx_of<T>::t C;
   const C *s[NRMAX];
   C *r[NRMAX];
   const C *u, *sw;
   const int *nbr;
   const unsigned char *bnd;
   int V, h, n;  // n: number of fields of this launch
   T mu, muo, coef;
   int masked;
   int xcdc;     // tiles per XCD run (xcd_tile; 0: one contiguous range per XCD)
};

template <typename T, int MU, bool FWD, int TPP, int NR>
__device__ inline void multi_phase(const MArgs<T> &a, int t, int p, int k, int e0, bool valid, const int *nb,
                                   typename cplx_of<T>::t (*S)[TPP * NR], typename cplx_of<T>::t (*X)[TPP * NR],
                                   const typename cplx_of<T>::t *sk, const typename cplx_of<T>::t *so,
                                   typename cplx_of<T>::t *acc_o, typename cplx_of<T>::t *acc_e)
{
   typedef typename cplx_of<T>::t C;
   const int e = e0 + p;
   if (valid)
   {
      int ny = nb[2 * MU + (FWD ? 0 : 1)];
      int y = ny & 0x7fffffff;
      int ly = y - e0;
      bool yin = (unsigned)ly < (unsigned)TPP;
      const C *up = link_ptr(a.u, e, FWD ? 2 * MU : 2 * MU + 1);
      C u[9];
#pragma unroll
      for (int q = 0; q < 9; q++) u[q] = up[q * SB];
      C h0[3], h1[3], c0[3], c1[3];
      if (!(a.masked && ny < 0))
      {
         C sy[12];
         if (yin)
         {
#pragma unroll
            for (int c = 0; c < 12; c++) sy[c] = S[c][ly * NR + k];
         }
         else
            load_spinor(sk, y, sy);
         spin_project<C, MU, FWD>(sy, h0, h1);
         su3_mul<C, !FWD>(u, h0, c0);
         su3_mul<C, !FWD>(u, h1, c1);
         spin_expand_add<C, MU, FWD>(c0, c1, acc_o);
      }
      if (yin)
      {
         spin_project<C, MU, !FWD>(so, h0, h1);
         su3_mul<C, FWD>(u, h0, c0);
         su3_mul<C, FWD>(u, h1, c1);
#pragma unroll
         for (int c = 0; c < 3; c++)
         {
            X[c][ly * NR + k] = c0[c];
            X[3 + c][ly * NR + k] = c1[c];
         }
      }
   }
   __syncthreads();
   if (valid)
   {
      int nz = nb[8 + 2 * MU + (FWD ? 1 : 0)];
      if (!(a.masked && nz < 0))
      {
         int z = nz & 0x7fffffff;
         if ((unsigned)(z - a.h - e0) < (unsigned)TPP)
         {
            C c0[3], c1[3];
#pragma unroll
            for (int c = 0; c < 3; c++)
            {
               c0[c] = X[c][t];
               c1[c] = X[3 + c][t];
            }
            spin_expand_add<C, MU, !FWD>(c0, c1, acc_e);
         }
         else
         {
            C sz[12];
            load_spinor(sk, z, sz);
            hop_one<C, MU, !FWD>(sz, link_ptr(a.u, z - a.h, FWD ? 2 * MU : 2 * MU + 1), acc_e);
         }
      }
   }
   __syncthreads();
}

template <typename T, int TPP, int NR>
__global__ __launch_bounds__(TPP * NR) void k_dw_multi(MArgs<T> a)
{
   typedef typename cplx_of<T>::t C;
   constexpr int NL = TPP * NR;
   __shared__ C S[12][NL];
   __shared__ C X[6][NL];
   const int t = threadIdx.x, p = t / NR, k = t % NR;
   const int e0 = xcd_tile(blockIdx.x, gridDim.x, a.xcdc) * TPP;
   const int e = e0 + p, o = a.h + e;
   const bool valid = e < a.h && k < a.n;
   const C *sk = a.s[k < a.n ? k : 0];
   C *rk = a.r[k < a.n ? k : 0];
   bool eb = false, ob = false;
   int nb[16];
   C so[12], acc_o[12], acc_e[12];
#pragma unroll
   for (int c = 0; c < 12; c++)
   {
      acc_o[c] = cmk<C>(0, 0);
      acc_e[c] = cmk<C>(0, 0);
      so[c] = cmk<C>(0, 0);
   }
   if (valid)
   {
#pragma unroll
      for (int q = 0; q < 8; q++)
      {
         nb[q] = a.nbr[(size_t)q * a.V + o];
         nb[8 + q] = a.nbr[(size_t)q * a.V + e];
      }
      eb = a.masked && a.bnd[e];
      ob = a.masked && a.bnd[o];
      C se[12];
      if (eb)
      {
#pragma unroll
         for (int c = 0; c < 12; c++) se[c] = cmk<C>(0, 0);
      }
      else
         load_spinor(sk, e, se);
#pragma unroll
      for (int c = 0; c < 12; c++) S[c][t] = se[c];
      if (!ob) load_spinor(sk, o, so);
   }
   __syncthreads();
   multi_phase<T, 0, true, TPP, NR>(a, t, p, k, e0, valid, nb, S, X, sk, so, acc_o, acc_e);
   multi_phase<T, 0, false, TPP, NR>(a, t, p, k, e0, valid, nb, S, X, sk, so, acc_o, acc_e);
   multi_phase<T, 1, true, TPP, NR>(a, t, p, k, e0, valid, nb, S, X, sk, so, acc_o, acc_e);
   multi_phase<T, 1, false, TPP, NR>(a, t, p, k, e0, valid, nb, S, X, sk, so, acc_o, acc_e);
   multi_phase<T, 2, true, TPP, NR>(a, t, p, k, e0, valid, nb, S, X, sk, so, acc_o, acc_e);
   multi_phase<T, 2, false, TPP, NR>(a, t, p, k, e0, valid, nb, S, X, sk, so, acc_o, acc_e);
   multi_phase<T, 3, true, TPP, NR>(a, t, p, k, e0, valid, nb, S, X, sk, so, acc_o, acc_e);
   multi_phase<T, 3, false, TPP, NR>(a, t, p, k, e0, valid, nb, S, X, sk, so, acc_o, acc_e);
   if (!valid) return;
   C out[12];
   if (ob)
      zero_spinor(rk, o);
   else
   {
      pauli_site<C, OQH_NT_SW>(a.sw, o, a.muo, so, out);
#pragma unroll
      for (int c = 0; c < 12; c++) out[c] = cadd(out[c], cscale(acc_o[c], a.coef));
      store_spinor(rk, o, out);
   }
   if (eb)
      zero_spinor(rk, e);
   else
   {
      C se[12];
#pragma unroll
      for (int c = 0; c < 12; c++) se[c] = S[c][t];
      pauli_site<C, OQH_NT_SW>(a.sw, e, a.mu, se, out);
#pragma unroll
      for (int c = 0; c < 12; c++) out[c] = cadd(out[c], cscale(acc_e[c], a.coef));
      store_spinor(rk, e, out);
   }
}

template <typename T>
__global__ __launch_bounds__(256) void k_zero_bnd(typename cplx_of<T>::t *s, const unsigned char *bnd,
                                                  int p0, int np)
{
   int k = blockIdx.x * 256 + threadIdx.x;
   if (k >= np) return;
   int x = p0 + k;
   if (bnd[x]) zero_spinor(s, x);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

#ifndef OQH_LINK_REC
#define OQH_LINK_REC 1
#endif
#ifndef OQH_SW_SFORM
#define OQH_SW_SFORM 1
#endif

// per odd point j and link l: compares row 2 with +-conj(row0 x row1); sets
// bit l of sign[j] for the minus sign and raises *dev to the largest
// deviation (as the bits of a non-negative double)
__global__ void k_link_classes(const double2 *__restrict__ ud, int h, unsigned char *sign, unsigned long long *dev)
{
   const int j = blockIdx.x * 256 + threadIdx.x;
   if (j >= h) return;
   unsigned char bits = 0;
   double mx = 0.0;
   for (int l = 0; l < 8; l++)
   {
      const double2 *up = link_ptr(ud, j, l);
      double2 u[9];
#pragma unroll
      for (int k = 0; k < 9; k++) u[k] = up[k * SB];
      double2 c[3] = {csub(cmul(u[1], u[5]), cmul(u[2], u[4])), csub(cmul(u[2], u[3]), cmul(u[0], u[5])),
                      csub(cmul(u[0], u[4]), cmul(u[1], u[3]))};
      double dp = 0.0, dm = 0.0;
#pragma unroll
      for (int k = 0; k < 3; k++)
      {
         // row 2 against conj(c) and -conj(c)
         dp = fmax(dp, fmax(fabs(u[6 + k].x - c[k].x), fabs(u[6 + k].y + c[k].y)));
         dm = fmax(dm, fmax(fabs(u[6 + k].x + c[k].x), fabs(u[6 + k].y - c[k].y)));
      }
      if (dm < dp) bits |= (unsigned char)(1u << l);
      mx = fmax(mx, fmin(dp, dm));
   }
   sign[j] = bits;
   atomicMax(dev, (unsigned long long)__double_as_longlong(mx));
}

// largest deviation of row 2 from +-conj(row0 x row1) that lets the
// operators rebuild row 2 (64 ulp; flowed fields stay near 1e-14 after 200
// Euler steps, unflowed random fields near 1e-15): the rebuilt operator then
// differs from the stored links by at most this much
constexpr double LINK_REC_TOL = 64.0 * DBL_EPSILON;

// link classes of the current FP64 field (once per field change): the
// operators rebuild row 2 of every link if all deviations are below
// LINK_REC_TOL
static void ensure_link_classes()
{
   Lib &L = lib();
   if (!OQH_LINK_REC || L.rec_version == L.ud_version) return;
   const int h = L.g.V / 2;
   static int nalloc = 0;
   if (!L.d_lsign || nalloc != h)
   {
      dev_free(L.d_lsign);
      L.d_lsign = (unsigned char *)dev_alloc((size_t)h + 1);
      nalloc = h;
   }
   unsigned long long *d_dev = (unsigned long long *)tmp_buffer(sizeof(unsigned long long));
   HIP_OK(hipMemsetAsync(d_dev, 0, sizeof(unsigned long long), L.stream));
   k_link_classes<<<(h + 255) / 256, 256, 0, L.stream>>>(L.d_ud, h, L.d_lsign, d_dev);
   unsigned long long bits = 0;
   HIP_OK(hipMemcpyAsync(&bits, d_dev, sizeof(bits), hipMemcpyDeviceToHost, L.stream));
   HIP_OK(hipStreamSynchronize(L.stream));
   double dev;
   memcpy(&dev, &bits, sizeof(dev));
   L.ud_rec = dev < LINK_REC_TOL;
   L.ud_rec_dev = dev;
   L.rec_version = L.ud_version;
}

template <typename T> static DArgs<T> base_args()
{
   Lib &L = lib();
   DArgs<T> a{};
   a.V = L.g.V;
   a.h = L.g.V / 2;
   a.nbr = L.d_nbr;
   a.bnd = L.d_bnd;
   a.masked = (L.bc != 3);
   a.u = (const typename cplx_of<T>::t *)(sizeof(T) == 8 ? (const void *)L.d_ud : (const void *)L.d_u);
   a.sw = (const typename cplx_of<T>::t *)(sizeof(T) == 8 ? (const void *)L.d_swd : (const void *)L.d_sw);
   // links rebuilt from two rows when the whole field allows it (the FP32
   // links are the rounded FP64 ones: same signs)
   a.lsign = (OQH_LINK_REC && L.ud_rec && L.rec_version == L.ud_version) ? L.d_lsign : nullptr;
   // SW term in the S-form (FP64, or its FP32 rounding) for the parities
   // where the field holds the uninverted term computed from the gauge field
   const bool swsok = OQH_SW_SFORM && L.sws_valid && (sizeof(T) == 8 || L.sws32_valid);
   a.sws_par = swsok ? ((L.swd_inv[0] ? 0 : 1) | (L.swd_inv[1] ? 0 : 2)) : 0;
   a.sws = a.sws_par ? (const typename cplx_of<T>::t *)(sizeof(T) == 8 ? (const void *)L.d_sws
                                                                         : (const void *)L.d_sws32)
                     : nullptr;
   a.x0 = L.d_x0;
   a.N0 = L.g.L[0];
   a.bc = L.bc;
   a.cx1 = (T)(4.0 + L.m0);
   a.cx3a = (T)(4.0 + L.m0 + L.cF[0] - 1.0);
   a.cx3b = (T)(4.0 + L.m0 + L.cF[1] - 1.0);
   // runs of as many tiles as there are thread blocks along x3, between 2
   // and 8 (measured: 32^3x64 with 8^4 thread blocks best at 4, 48^4 at 6;
   // the plain order and one contiguous range per XCD are 3-6 % slower)
   a.xcdc = std::max(2, std::min(8, L.g.L[3] / L.g.Ltrd[3]));
   return a;
}

template <typename T, int OP> static void launch(const DArgs<T> &a)
{
   Lib &L = lib();
#if OQH_DW_TILE
   if (OP == OP_DW)
   {
      constexpr int TP = TileSize<T>::n;
      unsigned nt = (unsigned)((a.h + TP - 1) / TP);
      if (!nt) return;
      const bool sws = a.sws_par == 3;
      if (a.lsign && sws)
         k_dw_tile<T, TP, true, true><<<nt, TP, 0, L.stream>>>(a);
      else if (a.lsign)
         k_dw_tile<T, TP, true, false><<<nt, TP, 0, L.stream>>>(a);
      else if (sws)
         k_dw_tile<T, TP, false, true><<<nt, TP, 0, L.stream>>>(a);
      else
         k_dw_tile<T, TP, false, false><<<nt, TP, 0, L.stream>>>(a);
      return;
   }
#endif
   unsigned nb = (OP == OP_DW) ? (unsigned)((a.h + 127) / 128) : (unsigned)((a.np + 255) / 256);
   if (nb == 0) return;
   k_dirac<T, OP><<<nb, 256, 0, L.stream>>>(a);
}

template <typename T> static void zero_bnd(typename cplx_of<T>::t *s, int p0, int np)
{
   Lib &L = lib();
   if (L.bc == 3 || np <= 0) return;
   k_zero_bnd<T><<<(np + 255) / 256, 256, 0, L.stream>>>(s, L.d_bnd, p0, np);
}

// state needed by an operator.  As in the reference, the SW field is used in
// whatever state sw_term() left it (the Schur complement expects the odd part
// inverted); it is computed here only if it was never computed for the
// current gauge field and parameters.
static bool prepare(const char *where, int prec)
{
   Lib &L = lib();
   if (!require_geometry(where)) return false;
   if (!L.ud_set)
   {
      fail(where, "gauge field is not set (oqh_import_ud)");
      return false;
   }
   if (!L.swd_valid) ensure_swd(0, 0);
   ensure_link_classes();
   if (prec == F32)
   {
      ensure_u();
      ensure_sw();
   }
   return true;
}

enum { K_DW, K_DEE, K_DOO, K_DOE, K_DEO, K_HAT };

template <typename T> static void apply(int kind, T mu, void *sp, void *rp, const char *where)
{
   typedef typename cplx_of<T>::t C;
   Lib &L = lib();
   int prec = sizeof(T) == 8 ? F64 : F32;
   if (!prepare(where, prec)) return;
   int V = L.g.V, h = V / 2;
   if ((kind == K_DW || kind == K_HAT) && sp == rp)
   {
      fail(where, "Input and output fields must be different");
      return;
   }
   bool rin = (kind == K_DEO);  // output is read (accumulated into)
   Res s = resolve(sp, prec, V, 0, true);
   Res r = resolve(rp, prec, V, 1, rin || kind != K_DW);
   if (!s.dev || !r.dev) return;
   if (s.off != 0 || r.off != 0)
   {
      fail(where, "the operators act on full fields (NSPIN points)");
      return;
   }
   DArgs<T> a = base_args<T>();
   a.s = (const C *)s.dev;
   a.r = (C *)r.dev;
   a.s_w = (C *)s.dev;
   T muo = L.eoflg ? (T)0 : mu;
   switch (kind)
   {
   case K_DW:
      a.mu = mu;
      a.muo = muo;
      a.coef = (T)-0.5;
      launch<T, OP_DW>(a);
      break;
   case K_DEE:
      a.mu = mu;
      a.p0 = 0;
      a.np = h;
      launch<T, OP_DIAG>(a);
      break;
   case K_DOO:
      a.mu = muo;
      a.p0 = h;
      a.np = h;
      launch<T, OP_DIAG>(a);
      break;
   case K_DOE:
      a.coef = (T)-0.5;
      a.p0 = h;
      a.np = h;
      launch<T, OP_DOE>(a);
      zero_bnd<T>((C *)s.dev, 0, h);   // cps(d)_int_bnd side effect
      break;
   case K_DEO:
      a.coef = (T)0.5;
      a.p0 = 0;
      a.np = h;
      launch<T, OP_DEO>(a);
      zero_bnd<T>((C *)s.dev, h, h);
      break;
   case K_HAT:
   {
      // odd intermediate t = A_oo^-1 (-1/2 hop(s_e)) in a scratch field
      a.t = (C *)tmp_buffer(units_for(V, 12) * sizeof(C));
#if OQH_HAT_TILE
      {
         constexpr int TPn = TileSize<T>::n;
         unsigned nt = (unsigned)((h + TPn - 1) / TPn);
         a.coef = (T)-0.5;
         if (a.lsign)
            k_hat1_tile<T, TileSize<T>::n, true><<<nt, TPn, 0, L.stream>>>(a);
         else
            k_hat1_tile<T, TileSize<T>::n, false><<<nt, TPn, 0, L.stream>>>(a);
         a.coef = (T)0.5;
         a.mu = mu;
         const bool sws = a.sws_par & 1;  // A_ee uninverted
         if (a.lsign && sws)
            k_hat2_tile<T, TileSize<T>::n, true, true><<<nt, TPn, 0, L.stream>>>(a);
         else if (a.lsign)
            k_hat2_tile<T, TileSize<T>::n, true, false><<<nt, TPn, 0, L.stream>>>(a);
         else if (sws)
            k_hat2_tile<T, TileSize<T>::n, false, true><<<nt, TPn, 0, L.stream>>>(a);
         else
            k_hat2_tile<T, TileSize<T>::n, false, false><<<nt, TPn, 0, L.stream>>>(a);
      }
#else
      a.coef = (T)-0.5;
      a.p0 = h;
      a.np = h;
      launch<T, OP_HAT1>(a);
      a.coef = (T)0.5;
      a.mu = mu;
      a.p0 = 0;
      a.np = h;
      launch<T, OP_HAT2>(a);
#endif
      zero_bnd<T>((C *)s.dev, 0, h);   // cps(d)_int_bnd side effect
      break;
   }
   }
   // inputs may have been zeroed on boundary points, as in the reference
   finish(s, prec, V, true);
   finish(r, prec, V, true);
}

}  // namespace oqh

using namespace oqh;

// zero a field on the boundary time slices of the current boundary
// conditions (time 0 and N0-1 for open bc), on the points of the set
template <typename T> static void bnd_zero_set(ptset_t set, void *f, int prec, const char *where)
{
   Lib &L = lib();
   if (!require_geometry(where)) return;
   const int V = L.g.V, h = V / 2;
   Res r = resolve(f, prec, V, 0, true);
   if (!r.dev) return;
   if (r.off)
   {
      fail(where, "the field must be a full field");
      return;
   }
   typedef typename cplx_of<T>::t C;
   if (set == ALL_PTS || set == EVEN_PTS) zero_bnd<T>((C *)r.dev, 0, h);
   if (set == ALL_PTS || set == ODD_PTS) zero_bnd<T>((C *)r.dev, h, h);
   finish(r, prec, V, true);
}

// lattice/bcnds.c:750 (bnd_sd2zero), bcnds.c:709 (bnd_s2zero)
extern "C" void bnd_sd2zero(ptset_t set, spinor_dble *sd)
{
   bnd_zero_set<double>(set, sd, F64, "bnd_sd2zero [bcnds.c]");
}

extern "C" void bnd_s2zero(ptset_t set, spinor *s)
{
   bnd_zero_set<float>(set, s, F32, "bnd_s2zero [bcnds.c]");
}

// whether the operators rebuild row 2 of the links of the current gauge
// field, and the largest deviation of a stored row 2 from the rebuilt one
extern "C" int oqh_link_rebuild(double *dev)
{
   Lib &L = lib();
   if (!L.ud_set)
   {
      fail("oqh_link_rebuild", "gauge field is not set");
      return 0;
   }
   ensure_link_classes();
   if (dev) *dev = L.ud_rec_dev;
   return L.ud_rec ? 1 : 0;
}

extern "C" void Dw_dble(double mu, spinor_dble *s, spinor_dble *r) { apply<double>(K_DW, mu, s, r, "Dw_dble [Dw_dble.c]"); }
extern "C" void Dwee_dble(double mu, spinor_dble *s, spinor_dble *r) { apply<double>(K_DEE, mu, s, r, "Dwee_dble [Dw_dble.c]"); }
extern "C" void Dwoo_dble(double mu, spinor_dble *s, spinor_dble *r) { apply<double>(K_DOO, mu, s, r, "Dwoo_dble [Dw_dble.c]"); }
extern "C" void Dwoe_dble(spinor_dble *s, spinor_dble *r) { apply<double>(K_DOE, 0.0, s, r, "Dwoe_dble [Dw_dble.c]"); }
extern "C" void Dweo_dble(spinor_dble *s, spinor_dble *r) { apply<double>(K_DEO, 0.0, s, r, "Dweo_dble [Dw_dble.c]"); }
extern "C" void Dwhat_dble(double mu, spinor_dble *s, spinor_dble *r) { apply<double>(K_HAT, mu, s, r, "Dwhat_dble [Dw_dble.c]"); }
extern "C" void Dw(float mu, spinor *s, spinor *r) { apply<float>(K_DW, mu, s, r, "Dw [Dw.c]"); }
extern "C" void Dwee(float mu, spinor *s, spinor *r) { apply<float>(K_DEE, mu, s, r, "Dwee [Dw.c]"); }
extern "C" void Dwoo(float mu, spinor *s, spinor *r) { apply<float>(K_DOO, mu, s, r, "Dwoo [Dw.c]"); }
extern "C" void Dwoe(spinor *s, spinor *r) { apply<float>(K_DOE, 0.0f, s, r, "Dwoe [Dw.c]"); }
extern "C" void Dweo(spinor *s, spinor *r) { apply<float>(K_DEO, 0.0f, s, r, "Dweo [Dw.c]"); }
extern "C" void Dwhat(float mu, spinor *s, spinor *r) { apply<float>(K_HAT, mu, s, r, "Dwhat [Dw.c]"); }

namespace oqh {

template <typename T, int NR> static void launch_multi(const MArgs<T> &a)
{
   constexpr int TPP = 256 / NR;
   unsigned nt = (unsigned)((a.h + TPP - 1) / TPP);
   if (nt) k_dw_multi<T, TPP, NR><<<nt, 256, 0, lib().stream>>>(a);
}

// n fields in chunks of up to NRMAX, each with the smallest power of two of
// lanes per pair that holds it
template <typename T> static void apply_multi(T mu, int n, void **sp, void **rp, const char *where)
{
   typedef typename cplx_of<T>::t C;
   Lib &L = lib();
   int prec = sizeof(T) == 8 ? F64 : F32;
   if (!prepare(where, prec)) return;
   if (n < 1 || !sp || !rp)
   {
      fail(where, "the number of fields must be positive");
      return;
   }
   MArgs<T> a{};
   // runs of 32 tiles per XCD (measured at 32^3x64: 0.517 ms per field
   // against 0.547 with one contiguous range per XCD; 2-128 within 2 %)
   a.xcdc = 32;
   a.u = (const C *)(sizeof(T) == 8 ? (const void *)L.d_ud : (const void *)L.d_u);
   a.sw = (const C *)(sizeof(T) == 8 ? (const void *)L.d_swd : (const void *)L.d_sw);
   a.nbr = L.d_nbr;
   a.bnd = L.d_bnd;
   a.V = L.g.V;
   a.h = L.g.V / 2;
   a.mu = mu;
   a.muo = L.eoflg ? (T)0 : mu;
   a.coef = (T)-0.5;
   a.masked = (L.bc != 3);
   for (int k0 = 0; k0 < n; k0 += NRMAX)
   {
      int m = (n - k0 < NRMAX) ? n - k0 : NRMAX;
      for (int k = 0; k < m; k++)
      {
         int o1 = 0, o2 = 0;
         void *s = device_field(sp[k0 + k], prec, &o1), *r = device_field(rp[k0 + k], prec, &o2);
         if (!s || !r || o1 || o2)
         {
            fail(where, "the fields must be full device fields of the library");
            return;
         }
         if (s == r)
         {
            fail(where, "Input and output fields must be different");
            return;
         }
         a.s[k] = (const C *)s;
         a.r[k] = (C *)r;
      }
      a.n = m;
      if (m == 1)
         launch_multi<T, 1>(a);
      else if (m == 2)
         launch_multi<T, 2>(a);
      else if (m <= 4)
         launch_multi<T, 4>(a);
      else
         launch_multi<T, 8>(a);
   }
}

}  // namespace oqh

extern "C" void oqh_Dw_dble_multi(double mu, int n, spinor_dble **s, spinor_dble **r)
{
   apply_multi<double>(mu, n, (void **)s, (void **)r, "oqh_Dw_dble_multi");
}

extern "C" void oqh_Dw_multi(float mu, int n, spinor **s, spinor **r)
{
   apply_multi<float>(mu, n, (void **)s, (void **)r, "oqh_Dw_multi");
}

extern "C" double oqh_time_dirac_multi(int prec, double mu, int n, void **s, void **r, int ncall)
{
   Lib &L = lib();
   if (ncall <= 0) return 0.0;
   if (prec) apply_multi<double>(mu, n, s, r, "oqh_time_dirac_multi");
   else apply_multi<float>((float)mu, n, s, r, "oqh_time_dirac_multi");
   HIP_OK(hipStreamSynchronize(L.stream));
   hipEvent_t e0, e1;
   HIP_OK(hipEventCreate(&e0));
   HIP_OK(hipEventCreate(&e1));
   HIP_OK(hipEventRecord(e0, L.stream));
   for (int k = 0; k < ncall; k++)
   {
      if (prec) apply_multi<double>(mu, n, s, r, "oqh_time_dirac_multi");
      else apply_multi<float>((float)mu, n, s, r, "oqh_time_dirac_multi");
   }
   HIP_OK(hipEventRecord(e1, L.stream));
   HIP_OK(hipEventSynchronize(e1));
   float ms = 0.0f;
   HIP_OK(hipEventElapsedTime(&ms, e0, e1));
   HIP_OK(hipEventDestroy(e0));
   HIP_OK(hipEventDestroy(e1));
   return (double)ms / ncall;
}

extern "C" double oqh_time_dirac(int op, double mu, void *s, void *r, int ncall)
{
   Lib &L = lib();
   if (ncall <= 0) return 0.0;
   // warm state (SW term, single-precision copies) outside the timed region
   if (op == 0) Dw_dble(mu, (spinor_dble *)s, (spinor_dble *)r);
   else if (op == 1) Dw((float)mu, (spinor *)s, (spinor *)r);
   else if (op == 2) Dwhat_dble(mu, (spinor_dble *)s, (spinor_dble *)r);
   else Dwhat((float)mu, (spinor *)s, (spinor *)r);
   HIP_OK(hipStreamSynchronize(L.stream));
   hipEvent_t e0, e1;
   HIP_OK(hipEventCreate(&e0));
   HIP_OK(hipEventCreate(&e1));
   HIP_OK(hipEventRecord(e0, L.stream));
   for (int k = 0; k < ncall; k++)
   {
      if (op == 0) Dw_dble(mu, (spinor_dble *)s, (spinor_dble *)r);
      else if (op == 1) Dw((float)mu, (spinor *)s, (spinor *)r);
      else if (op == 2) Dwhat_dble(mu, (spinor_dble *)s, (spinor_dble *)r);
      else Dwhat((float)mu, (spinor *)s, (spinor *)r);
   }
   HIP_OK(hipEventRecord(e1, L.stream));
   HIP_OK(hipEventSynchronize(e1));
   float ms = 0.0f;
   HIP_OK(hipEventElapsedTime(&ms, e0, e1));
   HIP_OK(hipEventDestroy(e0));
   HIP_OK(hipEventDestroy(e1));
   return (double)ms / ncall;
}

// ---------------------------------------------------------------------------
// apply_sw_dble, apply_sw (sw_term/pauli_dble.c:1840-1857, pauli.c:1107-1119):
// r = (A + i mu gamma5) s on vol points with the matrices m[2k], m[2k+1] of
// point k, for any array m of the reference's layout (for instance a copy
// from swdfld()/swfld() at an offset).  m is moved to the device in the
// site-block layout; s and r are device fields or host arrays and may
// coincide.
// ---------------------------------------------------------------------------

namespace oqh {

template <typename C>
__global__ __launch_bounds__(256) void k_apply_sw(const C *m, const C *s, C *r, int soff, int roff, int vol,
                                                  decltype(C::x) mu)
{
   const int k = blockIdx.x * 256 + threadIdx.x;
   if (k >= vol) return;
   C v[12], o[12];
   load_spinor(s, soff + k, v);
   pauli_site<C, false>(m, k, mu, v, o);
   store_spinor(r, roff + k, o);
}

template <typename C>
static void apply_sw_any(int vol, double mu, const void *m, void *s, void *r, int prec, const char *where)
{
   Lib &L = lib();
   if (!require_geometry(where) || vol <= 0) return;
   if (m == nullptr)
   {
      fail(where, "Attempt to access unallocated memory space");
      return;
   }
   const size_t eb = sizeof(C);
   const size_t aos_bytes = (size_t)vol * 36 * eb, blk_bytes = units_for(vol, 36) * eb;
   char *buf = (char *)tmp_buffer(aos_bytes + blk_bytes);
   if (!buf) return;
   C *aos = (C *)buf, *blk = (C *)(buf + aos_bytes);
   HIP_OK(hipMemcpyAsync(aos, m, aos_bytes, hipMemcpyHostToDevice, L.stream));
   if (prec == F64)
      aos_to_blocks_f64((const double2 *)aos, (double2 *)blk, vol, 36, L.stream);
   else
      aos_to_blocks_f32((const float2 *)aos, (float2 *)blk, vol, 36, L.stream);
   Res rs = resolve(s, prec, vol, 2, true);
   Res rr = resolve(r, prec, vol, 3, false);
   if (!rs.dev || !rr.dev) return;
   k_apply_sw<C><<<(unsigned)((vol + 255) / 256), 256, 0, L.stream>>>(blk, (const C *)rs.dev, (C *)rr.dev, rs.off,
                                                                        rr.off, vol, (decltype(C::x))mu);
   finish(rr, prec, vol, true);
   HIP_OK(hipStreamSynchronize(L.stream));
}

}  // namespace oqh

extern "C" void apply_sw_dble(int vol, double mu, pauli_dble *m, spinor_dble *s, spinor_dble *r)
{
   apply_sw_any<double2>(vol, mu, m, s, r, F64, "apply_sw_dble [pauli_dble.c]");
}

extern "C" void apply_sw(int vol, float mu, pauli *m, spinor *s, spinor *r)
{
   apply_sw_any<float2>(vol, mu, m, s, r, F32, "apply_sw [pauli.c]");
}
