// sap.hip -- Schwarz alternating procedure (SAP) on gfx950.
//
// Replaces sap/sap.c:592-790 (sap, update_flds0/1/2), sap/blk_solv.c:1117-1186
// (blk_mres, blk_eo_mres), dirac/Dw.c:1475-1907 (the block operators),
// dirac/Dw_bnd.c:801-1011 with sap/sap_com.c:465-501 (block boundary update)
// and the SAP_BLOCKS parts of block/blk_grid.c, map_u2blk.c, map_sw2blk.c and
// map_s2blk.c.
//
// One work-group handles one block of the current colour.  The block's three
// fields (solution s0, residual s1, work s2) are kept block-compact in a
// scratch area (site-block layout over the block-local point index, even
// points first, as the reference's block geometry, lattice/blk_geometry.c);
// links and SW matrices are read from the global single-precision fields
// through the block's embedding table imb.  All nmr minimal-residual steps of
// a block run inside the work-group: the block Dirac operator with Dirichlet
// exterior, the two block reductions (LDS tree in double precision, rounded
// to float as salg.c does) and the two vector updates, separated by
// work-group barriers.
//
// The reference computes the hopping terms from each block across its faces
// (Dw_bnd) and subtracts them from the residual of the neighbouring blocks
// after all blocks of the colour are done (sap_com).  Here the block
// corrections s0 are written to a global field and a second kernel gathers,
// for every face point of the blocks of the other colour, the hopping terms
// from its neighbours in other blocks, so that no point is updated by two
// work-items.
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <array>
#include <vector>

#include <hip/hip_fp16.h>

#include "dirac_dev.h"

namespace oqh {

constexpr int SWG = 256;      // work-group size of the block kernels
constexpr int SWG_MAX = 1024; // work-group size of the generic FP32 block solver for blocks of >= 1024 points
constexpr int NBF = 4;    // block fields of the API per block: s[0..2] + the Dwhat_blk intermediate
#ifndef OQH_SAP_REG
#define OQH_SAP_REG 1
#endif

struct SapGrid
{
   bool ready = false;
   int bs[4] = {0, 0, 0, 0};
   int nb = 0, nbh = 0, bvol = 0, bh = 0;
   size_t Ub = 0;                        // units per block field
   int *d_imb = nullptr;                 // [nb][bvol]
   int *d_bnb = nullptr;                 // [8][bvol], bvol = outside
   int *d_face[2] = {nullptr, nullptr};  // face points of colour-c blocks
   unsigned char *d_fmask[2] = {nullptr, nullptr};
   int nface[2] = {0, 0};
   float2 *d_swb = nullptr;              // FP32 SW field of the block solver
   int swb_set = -1;                     // ptset_t used for the inversion, -1 none
   int swb_epoch = -1;                   // SW field version it was made from
   float2 *d_scr = nullptr;              // [nbh][3][Ub]
   float2 *d_s0g = nullptr;              // global field of block corrections
   float2 *d_bs = nullptr;               // block fields of the API, [nb][3][Ub]
   double2 *d_bsd = nullptr;             // FP64 block fields of the API, [nb][NBF][Ub]
   bool need64 = false;                  // FP64 block operators in use: keep d_swb64
   __half2 *d_u16 = nullptr;             // FP16 links (FP16 variant of the solver)
   __half2 *d_swb16 = nullptr;           // FP16 block SW field
   float2 *d_s0gm = nullptr;             // block corrections of 16 sources
   double2 *d_swb64 = nullptr;           // FP64 block SW field (FP64 variant)
   int swb64_version = -1;
   double2 *d_h64[3] = {nullptr, nullptr, nullptr};  // FP64 psi, rho, s0g
   double2 *d_scr64 = nullptr;           // FP64 block fields [nbh][3][Ub]
   __half2 *d_h16[3] = {nullptr, nullptr, nullptr};  // FP16 psi, rho, s0g
   float *d_hsc[3] = {nullptr, nullptr, nullptr};    // their per-site scales
   int u16_epoch = -1, swb16_epoch = -1; // versions they were made from
   int swb_version = 0;                  // incremented by every assign_swb
   std::vector<int> imb;                 // host copy
   std::vector<int> bo;                  // [nb][4]
   int *d_ipp = nullptr;                 // face points of a block, faces 0..7 in turn
   int fofs[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // offsets of the faces in d_ipp
   float2 *d_bw = nullptr;               // Weyl fields of Dw_bnd, [nb][fofs[8]][6]
};

static SapGrid G;

template <typename C> struct SArgsT
{
   const C *u, *sw;
   const int *imb, *bnb, *nbr;
   const unsigned char *bnd;
   int V, h, bvol, bh, masked, eoflg, nmr;
   decltype(C::x) mu, muo;
   size_t Ub;
   C *psi, *rho, *s0g, *scr;
   int n0;             // first block of the colour
   // FP16 variant: psi, rho and s0g as FP16 spinors with per-site scales
   __half2 *psi16, *rho16, *s0g16;
   float *psc, *rsc, *s0sc;
   int pz;             // psi is zero on entry (not read)
   int pull;           // add the face terms from the other colour's s0g first
   // several right-hand sides in one launch (FP32): nsrc > 0 replaces psi,
   // rho, s0g by the fields of source blockIdx % nsrc (block kernel) or
   // blockIdx.y (face kernel)
   int nsrc;
   C *pl[16], *rl[16], *sl[16];
};
typedef SArgsT<float2> SArgs;   // FP32 and FP16 variants
typedef SArgsT<double2> SArgsD; // FP64 variant

// FP16 spinor with per-site scale: the 12 colour components of site x are
// stored as v / m in FP16 (same unit layout as the FP32 fields), m = the
// largest modulus of a real component of the spinor at x
__device__ inline void load_h16(const __half2 *h, const float *sc, int x, float2 *v)
{
   const float m = sc[x];
   const __half2 *p = h + (size_t)(x >> 6) * (12 * SB) + (x & (SB - 1));
#pragma unroll
   for (int c = 0; c < 12; c++)
   {
      float2 f = __half22float2(p[c * SB]);
      v[c] = make_float2(f.x * m, f.y * m);
   }
}

__device__ inline void store_h16(__half2 *h, float *sc, int x, const float2 *v)
{
   float m = 0.0f;
#pragma unroll
   for (int c = 0; c < 12; c++) m = fmaxf(m, fmaxf(fabsf(v[c].x), fabsf(v[c].y)));
   const float r = m > 0.0f ? 1.0f / m : 0.0f;
   sc[x] = m;
   __half2 *p = h + (size_t)(x >> 6) * (12 * SB) + (x & (SB - 1));
#pragma unroll
   for (int c = 0; c < 12; c++) p[c * SB] = __float22half2_rn(make_float2(v[c].x * r, v[c].y * r));
}

// link of direction mu, forward (x->x+mu) or backward (x-mu->x), seen from
// global point gx with global neighbour gy
template <typename C>
__device__ inline const C *blk_link(const SArgsT<C> &a, int gx, int gy, int mu, bool fwd)
{
   if (gx >= a.h) return link_ptr(a.u, gx - a.h, fwd ? 2 * mu : 2 * mu + 1);
   return link_ptr(a.u, gy - a.h, fwd ? 2 * mu + 1 : 2 * mu);
}

// acc += hopping terms from the in-block neighbours of block point i;
// neighbours on the global boundary (open bc) contribute nothing
template <int MU, typename C>
__device__ inline void blk_hop_mu(const SArgsT<C> &a, const int *imb, const C *S, int i, int gx, C *acc)
{
   C sp[12];
   int j = a.bnb[(2 * MU) * a.bvol + i];
   if (j < a.bvol)
   {
      int gy = imb[j];
      if (!(a.masked && a.bnd[gy]))
      {
         load_spinor(S, j, sp);
         hop_one<C, MU, true>(sp, blk_link(a, gx, gy, MU, true), acc);
      }
   }
   j = a.bnb[(2 * MU + 1) * a.bvol + i];
   if (j < a.bvol)
   {
      int gy = imb[j];
      if (!(a.masked && a.bnd[gy]))
      {
         load_spinor(S, j, sp);
         hop_one<C, MU, false>(sp, blk_link(a, gx, gy, MU, false), acc);
      }
   }
}

template <typename C>
__device__ inline void blk_hop(const SArgsT<C> &a, const int *imb, const C *S, int i, int gx, C *acc)
{
   blk_hop_mu<0>(a, imb, S, i, gx, acc);
   blk_hop_mu<1>(a, imb, S, i, gx, acc);
   blk_hop_mu<2>(a, imb, S, i, gx, acc);
   blk_hop_mu<3>(a, imb, S, i, gx, acc);
}

template <typename C> __device__ inline void zero12(C *v)
{
#pragma unroll
   for (int c = 0; c < 12; c++) v[c] = cmk<C>(0, 0);
}

// work-group sum of two doubles, broadcast to all threads
__device__ inline void wg_sum2(double &x, double &y, int nt)
{
   __shared__ double sh[2][SWG_MAX / 64];
#pragma unroll
   for (int m = 32; m >= 1; m >>= 1)
   {
      x += __shfl_xor(x, m);
      y += __shfl_xor(y, m);
   }
   int w = threadIdx.x >> 6;
   if ((threadIdx.x & 63) == 0)
   {
      sh[0][w] = x;
      sh[1][w] = y;
   }
   __syncthreads();
   x = 0.0;
   y = 0.0;
   for (int k = 0; k < (nt >> 6); k++)
   {
      x += sh[0][k];
      y += sh[1][k];
   }
   __syncthreads();
}

__device__ inline void wg_sum3(double &x, double &y, double &z, int nt)
{
   __shared__ double sh[3][SWG_MAX / 64];
#pragma unroll
   for (int m = 32; m >= 1; m >>= 1)
   {
      x += __shfl_xor(x, m);
      y += __shfl_xor(y, m);
      z += __shfl_xor(z, m);
   }
   int w = threadIdx.x >> 6;
   if ((threadIdx.x & 63) == 0)
   {
      sh[0][w] = x;
      sh[1][w] = y;
      sh[2][w] = z;
   }
   __syncthreads();
   x = y = z = 0.0;
   for (int k = 0; k < (nt >> 6); k++)
   {
      x += sh[0][k];
      y += sh[1][k];
      z += sh[2][k];
   }
   __syncthreads();
}

// S2 = Dw_blk S1 on the points [i0, i1) of the block (Dw.c:1475)
template <typename C>
__device__ void blk_dw(const SArgsT<C> &a, const int *imb, C *S1, C *S2, int i0, int i1, decltype(C::x) mu_e, int nt)
{
   typedef decltype(C::x) R;
   for (int i = i0 + threadIdx.x; i < i1; i += nt)
   {
      int gx = imb[i];
      C o[12];
      if (a.masked && a.bnd[gx])
      {
         zero12(o);
         store_spinor(S2, i, o);
         continue;
      }
      C acc[12], v[12];
      zero12(acc);
      blk_hop(a, imb, S1, i, gx, acc);
      load_spinor((const C *)S1, i, v);
      pauli_site(a.sw, gx, i < a.bh ? mu_e : a.muo, v, o);
#pragma unroll
      for (int c = 0; c < 12; c++) o[c] = cadd(o[c], cscale(acc[c], (R)-0.5));
      store_spinor(S2, i, o);
   }
}

// even points: S2_e = Dwhat_blk S1 (Dw.c:1828); the odd half of T holds the
// intermediate A_oo^-1 D_oe S1 (T = S2 inside the Schwarz cycles, a separate
// field for the block API, which leaves the odd half of S2 unchanged)
template <typename C>
__device__ void blk_dwhat(const SArgsT<C> &a, const int *imb, C *S1, C *S2, C *T, decltype(C::x) mu, int nt)
{
   typedef decltype(C::x) R;
   for (int i = a.bh + threadIdx.x; i < a.bvol; i += nt)
   {
      int gx = imb[i];
      C o[12];
      if (a.masked && a.bnd[gx])
      {
         zero12(o);
         store_spinor(T, i, o);
         continue;
      }
      C acc[12];
      zero12(acc);
      blk_hop(a, imb, S1, i, gx, acc);
#pragma unroll
      for (int c = 0; c < 12; c++) acc[c] = cscale(acc[c], (R)-0.5);
      pauli_site(a.sw, gx, (R)0, acc, o);
      store_spinor(T, i, o);
   }
   __syncthreads();
   for (int i = threadIdx.x; i < a.bh; i += nt)
   {
      int gx = imb[i];
      C o[12];
      if (a.masked && a.bnd[gx])
      {
         zero12(o);
         store_spinor(S2, i, o);
         continue;
      }
      C acc[12], v[12];
      zero12(acc);
      blk_hop(a, imb, T, i, gx, acc);
      load_spinor((const C *)S1, i, v);
      pauli_site(a.sw, gx, mu, v, o);
#pragma unroll
      for (int c = 0; c < 12; c++) o[c] = cadd(o[c], cscale(acc[c], (R)0.5));
      store_spinor(S2, i, o);
   }
}

// minimal residual steps on points [0, nv) (blk_solv.c:1117-1186); T: odd
// intermediate of Dwhat_blk (EO)
template <bool EO, typename C>
__device__ void blk_mres_dev(const SArgsT<C> &a, const int *imb, C *S0, C *S1, C *S2, C *T, decltype(C::x) mu, int nt)
{
   typedef decltype(C::x) R;
   const R rmin = sizeof(R) == 4 ? (R)FLT_MIN : (R)DBL_MIN;
   int nv = EO ? a.bh : a.bvol;
   for (int i = threadIdx.x; i < nv; i += nt) zero_spinor(S0, i);
   __syncthreads();
   for (int imr = 0; imr < a.nmr; imr++)
   {
      if (EO)
         blk_dwhat(a, imb, S1, S2, T, mu, nt);
      else
         blk_dw(a, imb, S1, S2, 0, a.bvol, mu, nt);
      __syncthreads();
      double rr = 0.0, zr = 0.0, zi = 0.0;
      for (int i = threadIdx.x; i < nv; i += nt)
      {
         const C *p2 = S2 + (size_t)(i >> 6) * (12 * SB) + (i & (SB - 1));
         const C *p1 = S1 + (size_t)(i >> 6) * (12 * SB) + (i & (SB - 1));
#pragma unroll
         for (int k = 0; k < 4; k++)
         {
            R nr = 0, pr = 0, pi = 0;
#pragma unroll
            for (int c = 0; c < 3; c++)
            {
               C u = p2[(3 * k + c) * SB], v = p1[(3 * k + c) * SB];
               nr += u.x * u.x + u.y * u.y;
               pr += u.x * v.x + u.y * v.y;
               pi += u.x * v.y - u.y * v.x;
            }
            rr += (double)nr;
            zr += (double)pr;
            zi += (double)pi;
         }
      }
      wg_sum3(rr, zr, zi, nt);
      R r = (R)rr;
      if (r < (R)2 * rmin) break;
      r = (R)1 / r;
      C z = cmk<C>((R)zr * r, (R)zi * r);
      for (int i = threadIdx.x; i < nv; i += nt)
      {
         C *p0 = S0 + (size_t)(i >> 6) * (12 * SB) + (i & (SB - 1));
         C *p1 = S1 + (size_t)(i >> 6) * (12 * SB) + (i & (SB - 1));
         const C *p2 = S2 + (size_t)(i >> 6) * (12 * SB) + (i & (SB - 1));
#pragma unroll
         for (int c = 0; c < 12; c++)
         {
            C s1 = p1[c * SB], s2 = p2[c * SB], s0 = p0[c * SB];
            p0[c * SB] = cmk<C>(s0.x + (z.x * s1.x - z.y * s1.y), s0.y + (z.x * s1.y + z.y * s1.x));
            p1[c * SB] = cmk<C>(s1.x - (z.x * s2.x - z.y * s2.y), s1.y - (z.x * s2.y + z.y * s2.x));
         }
      }
      __syncthreads();
   }
}

// isolv=1 preparation: S1_o = A_oo^-1 S1_o, then S1_e -= D_eo S1_o
// (Dwoo_blk, Dweo_blk of sap.c:759-761)
template <typename C> __device__ void blk_eo_prepare(const SArgsT<C> &a, const int *imb, C *S1, int nt)
{
   typedef decltype(C::x) R;
   for (int i = a.bh + threadIdx.x; i < a.bvol; i += nt)
   {
      int gx = imb[i];
      C v[12], o[12];
      if (a.masked && a.bnd[gx])
         zero12(o);
      else
      {
         load_spinor((const C *)S1, i, v);
         pauli_site(a.sw, gx, (R)0, v, o);
      }
      store_spinor(S1, i, o);
   }
   __syncthreads();
   for (int i = threadIdx.x; i < a.bh; i += nt)
   {
      int gx = imb[i];
      C o[12];
      if (a.masked && a.bnd[gx])
         zero12(o);
      else
      {
         C acc[12];
         zero12(acc);
         blk_hop(a, imb, S1, i, gx, acc);
         load_spinor((const C *)S1, i, o);
#pragma unroll
         for (int c = 0; c < 12; c++) o[c] = cadd(o[c], cscale(acc[c], (R)0.5));
      }
      store_spinor(S1, i, o);
   }
   __syncthreads();
}

// isolv=1 completion: odd solution (Dwoe_blk, Dwoo_blk; update_flds1/2)
template <typename C> __device__ void blk_eo_finish(const SArgsT<C> &a, const int *imb, C *S0, C *S1, int nt)
{
   typedef decltype(C::x) R;
   for (int i = a.bh + threadIdx.x; i < a.bvol; i += nt)
   {
      int gx = imb[i];
      C o[12];
      if (a.masked && a.bnd[gx])
         zero12(o);
      else
      {
         C acc[12];
         zero12(acc);
         blk_hop(a, imb, S0, i, gx, acc);
#pragma unroll
         for (int c = 0; c < 12; c++) acc[c] = cscale(acc[c], (R)-0.5);
         pauli_site(a.sw, gx, (R)0, acc, o);
      }
      // s0_o = s1_o - A_oo^-1 D_oe s0_e
      C r1[12];
      load_spinor((const C *)S1, i, r1);
#pragma unroll
      for (int c = 0; c < 12; c++) o[c] = csub(r1[c], o[c]);
      store_spinor(S0, i, o);
   }
}

// ---------------------------------------------------------------------------
// the Schwarz sweep over the blocks of one colour
// ---------------------------------------------------------------------------

template <int ISOLV, typename C, int TPB = SWG>
__global__ __launch_bounds__(TPB) void k_sap_blocks(SArgsT<C> a)
{
   typedef decltype(C::x) R;
   int n = a.n0 + blockIdx.x;
   const int *imb = a.imb + (size_t)n * a.bvol;
   C *S0 = a.scr + (size_t)blockIdx.x * 3 * a.Ub;
   C *S1 = S0 + a.Ub, *S2 = S1 + a.Ub;

   for (int i = threadIdx.x; i < a.bvol; i += TPB)
   {
      int gx = imb[i];
      C v[12];
      if (!ISOLV && a.masked && a.bnd[gx])
         zero12(v);
      else
         load_spinor((const C *)a.rho, gx, v);
      store_spinor(S1, i, v);
   }
   __syncthreads();

   if (ISOLV)
   {
      blk_eo_prepare(a, imb, S1, TPB);
      blk_mres_dev<true>(a, imb, S0, S1, S2, S2, a.mu, TPB);
      blk_eo_finish(a, imb, S0, S1, TPB);
      __syncthreads();
   }
   else
      blk_mres_dev<false>(a, imb, S0, S1, S2, S2, a.mu, TPB);

   for (int i = threadIdx.x; i < a.bvol; i += TPB)
   {
      int gx = imb[i];
      C s0[12], s1[12], p[12];
      load_spinor((const C *)S0, i, s0);
      load_spinor((const C *)S1, i, s1);
      load_spinor((const C *)a.psi, gx, p);
#pragma unroll
      for (int c = 0; c < 12; c++) p[c] = cadd(p[c], s0[c]);
      store_spinor(a.psi, gx, p);
      if (ISOLV && i >= a.bh)
      {
         // odd residual -i mu gamma5 s0 (update_flds1) or 0 (eoflg, update_flds2)
         R m = a.eoflg ? (R)0 : a.mu;
#pragma unroll
         for (int c = 0; c < 12; c++)
         {
            R cm = (c < 6) ? -m : m;
            s1[c] = cmk<C>(-cm * s0[c].y, cm * s0[c].x);
         }
      }
      store_spinor(a.rho, gx, s1);
      store_spinor(a.s0g, gx, s0);
   }
}

// ---------------------------------------------------------------------------
// the Schwarz sweep with the block held in registers (blocks of 256 points)
//
// Lane x of the work-group owns block point x: its Pauli pair, the links
// U(x,mu) to its in-block forward neighbours and the values of the block
// fields s0, s1, s2 at x stay in registers for the whole block solve, so the
// global memory is read once (links, SW, residual) and written once per
// block.  A hopping pass exchanges only spinors: each lane puts its input
// value into LDS (F), adds the forward terms of its neighbours x+mu from F and
// computes with its own link the backward term of x+mu, which it passes
// through LDS (X, two colour vectors after the spin projection).  The terms
// are added in the order of blk_hop (mu = 0..3, forward then backward), so
// the arithmetic is that of k_sap_blocks.
// ---------------------------------------------------------------------------

constexpr int RBV = 256;

// complex multiply-accumulate as FMA chains: acc + a b, acc + conj(a) b
__device__ inline float2 cmac(float2 acc, float2 a, float2 b)
{
   return make_float2(fmaf(a.x, b.x, fmaf(-a.y, b.y, acc.x)), fmaf(a.x, b.y, fmaf(a.y, b.x, acc.y)));
}
__device__ inline float2 cmacc(float2 acc, float2 a, float2 b)
{
   return make_float2(fmaf(a.x, b.x, fmaf(a.y, b.y, acc.x)), fmaf(a.x, b.y, fmaf(-a.y, b.x, acc.y)));
}

// c = U h or U^dag h (U in registers) with FMA chains
template <bool DAG> __device__ inline void su3_mul_f(const float2 *u, const float2 *h, float2 *c)
{
#pragma unroll
   for (int i = 0; i < 3; i++)
   {
      if (!DAG)
         c[i] = cmac(cmac(cmul(u[3 * i], h[0]), u[3 * i + 1], h[1]), u[3 * i + 2], h[2]);
      else
         c[i] = cmacc(cmacc(cmulc(u[i], h[0]), u[3 + i], h[1]), u[6 + i], h[2]);
   }
}

// storage of links and SW matrices in the block solver: float2, or __half2
// in the FP16 variant (arithmetic in FP32 either way)
__device__ inline float2 to_f2(float2 v) { return v; }
__device__ inline float2 to_f2(__half2 v) { return __half22float2(v); }
template <typename S> __device__ inline S from_f2(float2 v);
template <> __device__ inline float2 from_f2<float2>(float2 v) { return v; }
template <> __device__ inline __half2 from_f2<__half2>(float2 v) { return __float22half2_rn(v); }

template <typename S> struct RegBlock
{
   S u[4][9];  // U(x,mu) if x+mu is in the block
   S m[36];    // Pauli pair at x (units as in the SW field)
   int fwd[4], bwd[4];
   bool bp;    // open-bc boundary point
};

// out = (A + i mu) v on the two Weyl halves (pauli_half); the half hh = 0
// of A is read at m0 with stride st0, the half hh = 1 at m1 with stride st1
// (registers or LDS)
template <typename S>
__device__ inline void pauli_reg2(const S *m0, int st0, const S *m1, int st1, float mu, const float2 *v, float2 *out)
{
#pragma unroll
   for (int hh = 0; hh < 2; hh++)
   {
      float2 q[18];
      const S *m = hh ? m1 : m0;
      const int stride = hh ? st1 : st0;
#pragma unroll
      for (int k = 0; k < 18; k++) q[k] = to_f2(m[k * stride]);
      const float2 *w = v + 6 * hh;
      float2 *o = out + 6 * hh;
      float mm = hh ? -mu : mu;
      float d[6] = {q[0].x, q[0].y, q[1].x, q[1].y, q[2].x, q[2].y};
#pragma unroll
      for (int i = 0; i < 6; i++) o[i] = make_float2(d[i] * w[i].x - mm * w[i].y, d[i] * w[i].y + mm * w[i].x);
      int p = 3;
#pragma unroll
      for (int i = 0; i < 6; i++)
#pragma unroll
         for (int j = i + 1; j < 6; j++)
         {
            float2 hij = q[p];
            o[i] = cmac(o[i], hij, w[j]);
            o[j] = cmacc(o[j], hij, w[i]);
            p++;
         }
   }
}

template <typename S>
__device__ inline void pauli_reg(const S *m, float mu, const float2 *v, float2 *out, int stride = 1)
{
   pauli_reg2(m, stride, m + 18 * stride, stride, mu, v, out);
}

// OUT: parity of the points whose terms are needed (0 even, 1 odd, 2 all);
// lane x computes its own forward term if its output is needed and the
// backward term of x+mu if the output of x+mu is needed, so that for one
// parity the even and the odd lanes (separate wavefronts) share the work
template <int MU, int OUT, typename S>
__device__ inline void rb_hop_mu(const RegBlock<S> &b, bool odd, const float2 *v, float2 *acc, float2 (*F)[RBV],
                                 float2 (*X)[RBV])
{
   const int x = threadIdx.x;
   const int y = b.fwd[MU];
   const bool own = (OUT == 2) || (OUT == (int)odd);
   const bool other = (OUT == 2) || (OUT != (int)odd);
   if (y < RBV)
   {
      float2 h0[3], h1[3], c0[3], c1[3], u[9];
#pragma unroll
      for (int k = 0; k < 9; k++) u[k] = to_f2(b.u[MU][k]);
      if (own)
      {
         float2 sp[12];
#pragma unroll
         for (int c = 0; c < 12; c++) sp[c] = F[c][y];
         spin_project<float2, MU, true>(sp, h0, h1);
         su3_mul_f<false>(u, h0, c0);
         su3_mul_f<false>(u, h1, c1);
         spin_expand_add<float2, MU, true>(c0, c1, acc);
      }
      if (other)
      {
         // backward term of y from x through the same link
         spin_project<float2, MU, false>(v, h0, h1);
         su3_mul_f<true>(u, h0, c0);
         su3_mul_f<true>(u, h1, c1);
#pragma unroll
         for (int c = 0; c < 3; c++)
         {
            X[c][y] = c0[c];
            X[3 + c][y] = c1[c];
         }
      }
   }
   __syncthreads();
   if (own && b.bwd[MU] < RBV)
   {
      float2 c0[3], c1[3];
#pragma unroll
      for (int c = 0; c < 3; c++)
      {
         c0[c] = X[c][x];
         c1[c] = X[3 + c][x];
      }
      spin_expand_add<float2, MU, false>(c0, c1, acc);
   }
   __syncthreads();
}

// acc = sum of the hopping terms of the field v over in-block neighbours
// (at the points of parity OUT); values at boundary points of open bc count
// as zero (blk_hop skips them).  lb: LDS of 18 * RBV float2.
template <int OUT, typename S>
__device__ inline void rb_hop(const RegBlock<S> &b, bool odd, const float2 *vin, float2 *acc, float2 *lb)
{
   const int x = threadIdx.x;
   float2 v[12];
   if (OUT == 2)
   {
      float2(*F)[RBV] = (float2(*)[RBV])lb;
      float2(*X)[RBV] = (float2(*)[RBV])(lb + 12 * RBV);
#pragma unroll
      for (int c = 0; c < 12; c++)
      {
         v[c] = b.bp ? make_float2(0.0f, 0.0f) : vin[c];
         F[c][x] = v[c];
         acc[c] = make_float2(0.0f, 0.0f);
      }
      __syncthreads();
      rb_hop_mu<0, OUT>(b, odd, v, acc, F, X);
      rb_hop_mu<1, OUT>(b, odd, v, acc, F, X);
      rb_hop_mu<2, OUT>(b, odd, v, acc, F, X);
      rb_hop_mu<3, OUT>(b, odd, v, acc, F, X);
      return;
   }
   // one parity: the source lanes (the other parity) store their values
   // (F, indexed by the point within its parity), compute the backward
   // terms of all four directions with their own links and pass them on
   // (X[mu]), while the output lanes compute their forward terms from F;
   // two barriers per pass.  The output lanes add the forward terms of
   // mu = 0..3, then the backward ones.
   constexpr int RBH = RBV / 2;
   float2(*F)[RBH] = (float2(*)[RBH])lb;
   float2(*X)[6][RBH] = (float2(*)[6][RBH])(lb + 12 * RBH);
   const bool own = (OUT == (int)odd);
#pragma unroll
   for (int c = 0; c < 12; c++)
   {
      v[c] = b.bp ? make_float2(0.0f, 0.0f) : vin[c];
      if (!own) F[c][x & (RBH - 1)] = v[c];
      acc[c] = make_float2(0.0f, 0.0f);
   }
   __syncthreads();
#pragma unroll
   for (int mu = 0; mu < 4; mu++)
   {
      const int y = b.fwd[mu];
      if (y < RBV)
      {
         float2 h0[3], h1[3], c0[3], c1[3], u[9];
#pragma unroll
         for (int k = 0; k < 9; k++) u[k] = to_f2(b.u[mu][k]);
         if (own)
         {
            float2 sp[12];
#pragma unroll
            for (int c = 0; c < 12; c++) sp[c] = F[c][y & (RBH - 1)];
            if (mu == 0) { spin_project<float2, 0, true>(sp, h0, h1); }
            if (mu == 1) { spin_project<float2, 1, true>(sp, h0, h1); }
            if (mu == 2) { spin_project<float2, 2, true>(sp, h0, h1); }
            if (mu == 3) { spin_project<float2, 3, true>(sp, h0, h1); }
            su3_mul_f<false>(u, h0, c0);
            su3_mul_f<false>(u, h1, c1);
            if (mu == 0) spin_expand_add<float2, 0, true>(c0, c1, acc);
            if (mu == 1) spin_expand_add<float2, 1, true>(c0, c1, acc);
            if (mu == 2) spin_expand_add<float2, 2, true>(c0, c1, acc);
            if (mu == 3) spin_expand_add<float2, 3, true>(c0, c1, acc);
         }
         else
         {
            if (mu == 0) spin_project<float2, 0, false>(v, h0, h1);
            if (mu == 1) spin_project<float2, 1, false>(v, h0, h1);
            if (mu == 2) spin_project<float2, 2, false>(v, h0, h1);
            if (mu == 3) spin_project<float2, 3, false>(v, h0, h1);
            su3_mul_f<true>(u, h0, c0);
            su3_mul_f<true>(u, h1, c1);
#pragma unroll
            for (int c = 0; c < 3; c++)
            {
               X[mu][c][y & (RBH - 1)] = c0[c];
               X[mu][3 + c][y & (RBH - 1)] = c1[c];
            }
         }
      }
   }
   __syncthreads();
   if (own)
   {
#pragma unroll
      for (int mu = 0; mu < 4; mu++)
      {
         if (b.bwd[mu] < RBV)
         {
            float2 c0[3], c1[3];
#pragma unroll
            for (int c = 0; c < 3; c++)
            {
               c0[c] = X[mu][c][x & (RBH - 1)];
               c1[c] = X[mu][3 + c][x & (RBH - 1)];
            }
            if (mu == 0) spin_expand_add<float2, 0, false>(c0, c1, acc);
            if (mu == 1) spin_expand_add<float2, 1, false>(c0, c1, acc);
            if (mu == 2) spin_expand_add<float2, 2, false>(c0, c1, acc);
            if (mu == 3) spin_expand_add<float2, 3, false>(c0, c1, acc);
         }
      }
   }
}

// block sums of |s2|^2 and (s2, s1) as in blk_mres_dev, lanes with act
__device__ inline void rb_sums(bool act, const float2 *s2, const float2 *s1, double &rr, double &zr, double &zi)
{
   rr = zr = zi = 0.0;
   if (act)
   {
#pragma unroll
      for (int k = 0; k < 4; k++)
      {
         float nr = 0.0f, pr = 0.0f, pi = 0.0f;
#pragma unroll
         for (int c = 0; c < 3; c++)
         {
            float2 u = s2[3 * k + c], v = s1[3 * k + c];
            nr += u.x * u.x + u.y * u.y;
            pr += u.x * v.x + u.y * v.y;
            pi += u.x * v.y - u.y * v.x;
         }
         rr += (double)nr;
         zr += (double)pr;
         zi += (double)pi;
      }
   }
   wg_sum3(rr, zr, zi, RBV);
}

// rho(x) += 1/2 hopping terms into block point x from the corrections s0g
// of the neighbouring blocks (the other colour), as k_sap_faces applies
// them, taken here at the start of the block solve
#ifndef OQH_PULL_GROUP
#define OQH_PULL_GROUP 2
#endif
template <typename S>
__device__ inline void pull_faces(const SArgs &a, const RegBlock<S> &b, const S *ub, int gx, float2 *s1)
{
   constexpr bool H = sizeof(S) == 4;
   const bool odd = gx >= a.h;
   float2 acc[12], h0[3], h1[3], c0[3], c1[3];
   zero12(acc);
   // as in k_sap_faces: one neighbour and one link load per direction,
   // OQH_PULL_GROUP directions ahead of the arithmetic
   constexpr int FG = OQH_PULL_GROUP;
   for (int m0 = 0; m0 < 4; m0 += FG)
   {
      float2 sp[FG][12], u[FG][9];
#pragma unroll
      for (int j = 0; j < FG; j++)
      {
         const int MU = m0 + j;
         const bool fw = b.fwd[MU] == RBV, bw = b.bwd[MU] == RBV;
         if (fw || bw)
         {
            const int y = a.nbr[(size_t)(fw ? 2 * MU : 2 * MU + 1) * a.V + gx] & 0x7fffffff;
            if (H)
               load_h16(a.s0g16, a.s0sc, y, sp[j]);
            else
               load_spinor((const float2 *)a.s0g, y, sp[j]);
            const int J = odd ? gx - a.h : y - a.h;
            const int l = (odd == fw) ? 2 * MU : 2 * MU + 1;
            const S *up = ub + (J >> 6) * (size_t)(72 * SB) + (J & (SB - 1)) + (size_t)(9 * l) * SB;
#pragma unroll
            for (int q = 0; q < 9; q++) u[j][q] = to_f2(up[q * SB]);
         }
      }
#pragma unroll
      for (int j = 0; j < FG; j++)
      {
         const int MU = m0 + j;
         const bool fw = b.fwd[MU] == RBV, bw = b.bwd[MU] == RBV;
#define OQH_PULL_TERM(M)                                                               \
   if (MU == M)                                                                       \
   {                                                                                  \
      if (fw)                                                                         \
      {                                                                               \
         spin_project<float2, M, true>(sp[j], h0, h1);                                \
         su3_mul_f<false>(u[j], h0, c0);                                              \
         su3_mul_f<false>(u[j], h1, c1);                                              \
         spin_expand_add<float2, M, true>(c0, c1, acc);                               \
      }                                                                               \
      else if (bw)                                                                    \
      {                                                                               \
         spin_project<float2, M, false>(sp[j], h0, h1);                               \
         su3_mul_f<true>(u[j], h0, c0);                                               \
         su3_mul_f<true>(u[j], h1, c1);                                               \
         spin_expand_add<float2, M, false>(c0, c1, acc);                              \
      }                                                                               \
   }
         OQH_PULL_TERM(0)
         OQH_PULL_TERM(1)
         OQH_PULL_TERM(2)
         OQH_PULL_TERM(3)
#undef OQH_PULL_TERM
      }
   }
#pragma unroll
   for (int c = 0; c < 12; c++) s1[c] = cadd(s1[c], cscale(acc[c], 0.5f));
}

#ifndef OQH_SAP_WAVES16
#define OQH_SAP_WAVES16 2
#endif
#ifndef OQH_SAP_WAVES32
#define OQH_SAP_WAVES32 2
#endif
#ifndef OQH_SAP_MHALF
#define OQH_SAP_MHALF 1
#endif

// Two wavefronts per SIMD (256 registers per lane and at most 80 KiB of LDS
// per work-group): the FP16 variant keeps its links in FP16 registers and
// its Pauli matrices in LDS; the FP32 variant keeps the first Pauli matrix
// of each point in LDS and the second in registers (OQH_SAP_MHALF; with
// both in registers it needs 310 registers and runs one wavefront per SIMD,
// 1.50 against 0.82 ms per colour at 48^4).
// MASKED: open-bc boundary points exist (bc != 3); without them the masking
// of boundary values is compiled out
template <int ISOLV, typename S, bool MASKED>
__global__ __launch_bounds__(RBV)
__attribute__((amdgpu_waves_per_eu(sizeof(S) == 4 ? OQH_SAP_WAVES16 : OQH_SAP_WAVES32)))
void k_sap_reg(SArgs a0, const S *u16, const S *sw16)
{
   __shared__ float2 LB[18 * RBV];  // hopping exchange (rb_hop)
   // FP16 variant: the Pauli matrices of the block live in LDS, and psi,
   // rho, s0g are FP16 spinors with per-site scales; FP32 variant with
   // OQH_SAP_MHALF: the first Pauli matrix of each point in LDS, the second
   // in registers
   constexpr bool MLDS = sizeof(S) == 4;
   constexpr bool MHALF = !MLDS && OQH_SAP_MHALF;
   constexpr bool MALL = MLDS;
   __shared__ S MS[MALL ? 36 : (MHALF ? 18 : 1)][RBV];
   // several sources: consecutive work-groups take the same block for
   // different sources, so its links and SW matrices come from the cache
   const int src = a0.nsrc ? (int)(blockIdx.x % a0.nsrc) : 0;
   const int bi = a0.nsrc ? (int)(blockIdx.x / a0.nsrc) : (int)blockIdx.x;
   SArgs a = a0;
   if (a0.nsrc)
   {
      a.psi = a0.pl[src];
      a.rho = a0.rl[src];
      a.s0g = a0.sl[src];
   }
   const int x = threadIdx.x;
   const int n = a.n0 + bi;
   const int *imb = a.imb + (size_t)n * RBV;
   const int gx = imb[x];
   const bool even = x < a.bh;
   RegBlock<S> b;
   b.bp = MASKED && a.bnd[gx];
   // links and SW matrices: the FP32 fields, or their FP16 copies
   const S *ub = sizeof(S) == sizeof(float2) ? (const S *)a.u : u16;
   const S *swb = sizeof(S) == sizeof(float2) ? (const S *)a.sw : sw16;
#pragma unroll
   for (int mu = 0; mu < 4; mu++)
   {
      b.fwd[mu] = a.bnb[(2 * mu) * RBV + x];
      b.bwd[mu] = a.bnb[(2 * mu + 1) * RBV + x];
      if (b.fwd[mu] < RBV)
      {
         const S *up = ub + (blk_link(a, gx, imb[b.fwd[mu]], mu, true) - a.u);
#pragma unroll
         for (int k = 0; k < 9; k++) b.u[mu][k] = up[k * SB];
      }
   }
   {
      const S *m = swb + (size_t)(gx >> 6) * (36 * SB) + (gx & (SB - 1));
#pragma unroll
      for (int k = 0; k < 36; k++)
      {
         if (MALL || (MHALF && k < 18))
            MS[k][x] = m[k * SB];
         else
            b.m[k] = m[k * SB];
      }
   }
   const S *mp0 = (MALL || MHALF) ? &MS[0][x] : b.m;
   const int mst0 = (MALL || MHALF) ? RBV : 1;
   const S *mp1 = MALL ? &MS[18][x] : b.m + 18;
   const int mst1 = MALL ? RBV : 1;
   float2 s0[12], s1[12], s2[12], acc[12];
   if (MLDS)
      load_h16(a.rho16, a.rsc, gx, s1);
   else
      load_spinor((const float2 *)a.rho, gx, s1);
   // only the FP16 variant pulls the face terms (the FP32 cycles use the
   // face kernel); the pull is not compiled into the FP32 kernel, whose
   // registers it would crowd
   if constexpr (MLDS)
      if (a.pull) pull_faces<S>(a, b, ub, gx, s1);
#pragma unroll
   for (int c = 0; c < 12; c++)
   {
      s0[c] = make_float2(0.0f, 0.0f);
      s2[c] = make_float2(0.0f, 0.0f);
      if (!ISOLV && b.bp) s1[c] = make_float2(0.0f, 0.0f);
   }

   if (ISOLV)
   {
      // odd points: s1 = A_oo^-1 s1, then even points: s1 += 1/2 hops
      // (Dwoo_blk, Dweo_blk of sap.c:759-761)
      if (!even)
      {
         float2 o[12];
         pauli_reg2(mp0, mst0, mp1, mst1, 0.0f, s1, o);
#pragma unroll
         for (int c = 0; c < 12; c++) s1[c] = b.bp ? make_float2(0.0f, 0.0f) : o[c];
      }
      rb_hop<0>(b, !even, s1, acc, LB);
      if (even)
      {
#pragma unroll
         for (int c = 0; c < 12; c++)
            s1[c] = b.bp ? make_float2(0.0f, 0.0f) : cadd(s1[c], cscale(acc[c], 0.5f));
      }
   }

   // minimal residual steps (blk_solv.c:1117-1186)
   for (int imr = 0; imr < a.nmr; imr++)
   {
      if (ISOLV)
      {
         // s2 (odd) = A_oo^-1 (-1/2 hops of s1_e); s2 (even) = A_ee s1 + 1/2 hops of s2_o
         rb_hop<1>(b, !even, s1, acc, LB);
         if (!even)
         {
#pragma unroll
            for (int c = 0; c < 12; c++) acc[c] = cscale(acc[c], -0.5f);
            pauli_reg2(mp0, mst0, mp1, mst1, 0.0f, acc, s2);
            if (b.bp)
            {
#pragma unroll
               for (int c = 0; c < 12; c++) s2[c] = make_float2(0.0f, 0.0f);
            }
         }
         rb_hop<0>(b, !even, s2, acc, LB);
         if (even)
         {
            pauli_reg2(mp0, mst0, mp1, mst1, a.mu, s1, s2);
#pragma unroll
            for (int c = 0; c < 12; c++)
               s2[c] = b.bp ? make_float2(0.0f, 0.0f) : cadd(s2[c], cscale(acc[c], 0.5f));
         }
      }
      else
      {
         rb_hop<2>(b, !even, s1, acc, LB);
         pauli_reg2(mp0, mst0, mp1, mst1, even ? a.mu : a.muo, s1, s2);
#pragma unroll
         for (int c = 0; c < 12; c++)
            s2[c] = b.bp ? make_float2(0.0f, 0.0f) : cadd(s2[c], cscale(acc[c], -0.5f));
      }
      double rr, zr, zi;
      rb_sums(ISOLV ? even : true, s2, s1, rr, zr, zi);
      float r = (float)rr;
      if (r < 2.0f * FLT_MIN) break;
      r = 1.0f / r;
      float2 z = make_float2((float)zr * r, (float)zi * r), mz = make_float2(-z.x, -z.y);
      if (!ISOLV || even)
      {
#pragma unroll
         for (int c = 0; c < 12; c++)
         {
            s0[c] = cmac(s0[c], z, s1[c]);
            s1[c] = cmac(s1[c], mz, s2[c]);
         }
      }
   }

   if (ISOLV)
   {
      // odd solution s0_o = s1_o - A_oo^-1 (-1/2 hops of s0_e) (Dwoe_blk, Dwoo_blk)
      rb_hop<1>(b, !even, s0, acc, LB);
      if (!even)
      {
         float2 o[12];
#pragma unroll
         for (int c = 0; c < 12; c++) acc[c] = cscale(acc[c], -0.5f);
         pauli_reg2(mp0, mst0, mp1, mst1, 0.0f, acc, o);
#pragma unroll
         for (int c = 0; c < 12; c++)
         {
            if (b.bp) o[c] = make_float2(0.0f, 0.0f);
            s0[c] = csub(s1[c], o[c]);
         }
      }
   }

   float2 p[12];
   if (a.pz)
      zero12(p);
   else if (MLDS)
      load_h16(a.psi16, a.psc, gx, p);
   else
      load_spinor((const float2 *)a.psi, gx, p);
#pragma unroll
   for (int c = 0; c < 12; c++) p[c] = cadd(p[c], s0[c]);
   if (MLDS)
      store_h16(a.psi16, a.psc, gx, p);
   else
      store_spinor(a.psi, gx, p);
   if (ISOLV && !even)
   {
      // odd residual -i mu gamma5 s0 (update_flds1) or 0 (eoflg, update_flds2)
      float m = a.eoflg ? 0.0f : a.mu;
#pragma unroll
      for (int c = 0; c < 12; c++)
      {
         float cm = (c < 6) ? -m : m;
         s1[c] = make_float2(-cm * s0[c].y, cm * s0[c].x);
      }
   }
   if (MLDS)
   {
      store_h16(a.rho16, a.rsc, gx, s1);
      store_h16(a.s0g16, a.s0sc, gx, s0);
   }
   else
   {
      store_spinor(a.rho, gx, s1);
      store_spinor(a.s0g, gx, s0);
   }
}

#ifndef OQH_FACE_GROUP
#define OQH_FACE_GROUP 4
#endif

// rho(y) -= hopping terms into y from the block corrections of the
// neighbouring blocks, for the face points y of the blocks of one colour
// (S = __half2: FP16 links, FP16 s0g and rho with per-site scales)
template <typename S>
__global__ __launch_bounds__(256) void k_sap_faces(SArgs a0, const int *face, const unsigned char *fmask, int nf,
                                                   const S *u16)
{
   constexpr bool H = sizeof(S) == 4;
   SArgs a = a0;
   if (a0.nsrc)
   {
      a.rho = a0.rl[blockIdx.y];
      a.s0g = a0.sl[blockIdx.y];
   }
   int k = blockIdx.x * 256 + threadIdx.x;
   if (k >= nf) return;
   int y = face[k];
   int m = fmask[k];
   float2 acc[12], h0[3], h1[3], c0[3], c1[3];
   zero12(acc);
   int odd = y >= a.h;
   const S *ub = H ? u16 : (const S *)a.u;
   // a point of a block of width >= 2 lies on at most one of the faces -mu,
   // +mu: the loads of each direction are issued for all lanes at once (one
   // neighbour and one link per direction and lane), OQH_FACE_GROUP
   // directions ahead of the arithmetic, which adds the terms in the order
   // mu = 0..3 (forward before backward)
   constexpr int FG = OQH_FACE_GROUP;
   for (int m0 = 0; m0 < 4; m0 += FG)
   {
      float2 sp[FG][12], u[FG][9];
#pragma unroll
      for (int j = 0; j < FG; j++)
      {
         const int MU = m0 + j;
         const bool fw = m & (1 << (2 * MU)), bw = m & (2 << (2 * MU));
         if (fw || bw)
         {
            const int dir = fw ? 2 * MU : 2 * MU + 1;
            const int x = a.nbr[(size_t)dir * a.V + y] & 0x7fffffff;
            if (H)
               load_h16(a.s0g16, a.s0sc, x, sp[j]);
            else
               load_spinor((const float2 *)a.s0g, x, sp[j]);
            // forward: U(y, mu) (odd y) or U(x, mu) stored at the odd x
            const int J = odd ? y - a.h : x - a.h;
            const int l = (odd == fw) ? 2 * MU : 2 * MU + 1;
            const S *up = ub + (J >> 6) * (size_t)(72 * SB) + (J & (SB - 1)) + (size_t)(9 * l) * SB;
#pragma unroll
            for (int q = 0; q < 9; q++) u[j][q] = to_f2(up[q * SB]);
         }
      }
#pragma unroll
      for (int j = 0; j < FG; j++)
      {
         const int MU = m0 + j;
         const bool fw = m & (1 << (2 * MU)), bw = m & (2 << (2 * MU));
#define OQH_FACE_TERM(M)                                                               \
   if (MU == M)                                                                       \
   {                                                                                  \
      if (fw)                                                                         \
      {                                                                               \
         spin_project<float2, M, true>(sp[j], h0, h1);                                \
         su3_mul_f<false>(u[j], h0, c0);                                              \
         su3_mul_f<false>(u[j], h1, c1);                                              \
         spin_expand_add<float2, M, true>(c0, c1, acc);                               \
      }                                                                               \
      else if (bw)                                                                    \
      {                                                                               \
         spin_project<float2, M, false>(sp[j], h0, h1);                               \
         su3_mul_f<true>(u[j], h0, c0);                                               \
         su3_mul_f<true>(u[j], h1, c1);                                               \
         spin_expand_add<float2, M, false>(c0, c1, acc);                              \
      }                                                                               \
   }
         OQH_FACE_TERM(0)
         OQH_FACE_TERM(1)
         OQH_FACE_TERM(2)
         OQH_FACE_TERM(3)
#undef OQH_FACE_TERM
      }
   }
   float2 r[12];
   if (H)
      load_h16(a.rho16, a.rsc, y, r);
   else
      load_spinor((const float2 *)a.rho, y, r);
#pragma unroll
   for (int c = 0; c < 12; c++) r[c] = cadd(r[c], cscale(acc[c], 0.5f));
   if (H)
      store_h16(a.rho16, a.rsc, y, r);
   else
      store_spinor(a.rho, y, r);
}

// ---------------------------------------------------------------------------
// block operators on the block fields of the API (one work-group per call)
// ---------------------------------------------------------------------------

enum BlkOp { BOP_DW = 0, BOP_DWHAT, BOP_MRES, BOP_EO_MRES, BOP_DWEE, BOP_DWOO, BOP_DWOE, BOP_DWEO };

template <typename C>
__global__ __launch_bounds__(SWG) void k_blk_op(SArgsT<C> a, int n, int op, int k, int l, decltype(C::x) mu)
{
   typedef decltype(C::x) R;
   const int *imb = a.imb + (size_t)n * a.bvol;
   C *F = a.scr + (size_t)n * NBF * a.Ub;  // the block's fields 0,1,2
   C *Sk = F + (size_t)k * a.Ub, *Sl = F + (size_t)l * a.Ub;
   C *T = F + 3 * a.Ub;  // odd intermediate of Dwhat_blk
   // open bc: the operators zero their input on the boundary points (all of
   // them for Dw_blk, the even ones for Dwhat_blk, Dwee_blk and Dwoe_blk,
   // the odd ones for Dwoo_blk and Dweo_blk); in the MinRes solvers the
   // input is field 1
   if (a.masked)
   {
      const bool even = (op == BOP_DWHAT || op == BOP_EO_MRES || op == BOP_DWEE || op == BOP_DWOE);
      const bool odd = (op == BOP_DWOO || op == BOP_DWEO);
      C *In = (op == BOP_MRES || op == BOP_EO_MRES) ? F + a.Ub : Sk;
      for (int i = threadIdx.x; i < a.bvol; i += SWG)
         if (a.bnd[imb[i]] && (!even || i < a.bh) && (!odd || i >= a.bh)) zero_spinor(In, i);
      __syncthreads();
   }
   if (op == BOP_DW)
      blk_dw(a, imb, Sk, Sl, 0, a.bvol, mu, SWG);
   else if (op == BOP_DWHAT)
   {
      // Dwhat_blk leaves the odd half of the output unchanged
      blk_dwhat(a, imb, Sk, Sl, T, mu, SWG);
   }
   else if (op == BOP_DWEE || op == BOP_DWOO)
   {
      // (A + i mu gamma5) on one parity (Dw.c:1584-1694); the fields may coincide
      const bool ee = op == BOP_DWEE;
      const R m = ee ? mu : a.muo;
      for (int i = (ee ? 0 : a.bh) + threadIdx.x; i < (ee ? a.bh : a.bvol); i += SWG)
      {
         const int gx = imb[i];
         C v[12], o[12];
         if (a.masked && a.bnd[gx])
            zero12(o);
         else
         {
            load_spinor((const C *)Sk, i, v);
            pauli_site(a.sw, gx, m, v, o);
         }
         store_spinor(Sl, i, o);
      }
   }
   else if (op == BOP_DWOE || op == BOP_DWEO)
   {
      // hopping terms between the parities with zero exterior (Dw.c:1696-1826):
      // Dwoe_blk assigns D_oe s to the odd points, Dweo_blk subtracts D_eo s
      // from the even points; the output parity is never read
      const bool oe = op == BOP_DWOE;
      for (int i = (oe ? a.bh : 0) + threadIdx.x; i < (oe ? a.bvol : a.bh); i += SWG)
      {
         const int gx = imb[i];
         C acc[12], o[12];
         zero12(acc);
         if (a.masked && a.bnd[gx])
            zero12(o);
         else
         {
            blk_hop(a, imb, (const C *)Sk, i, gx, acc);
            if (oe)
            {
#pragma unroll
               for (int c = 0; c < 12; c++) o[c] = cscale(acc[c], (R)-0.5);
            }
            else
            {
               load_spinor((const C *)Sl, i, o);
#pragma unroll
               for (int c = 0; c < 12; c++) o[c] = cadd(o[c], cscale(acc[c], (R)0.5));
            }
         }
         store_spinor(Sl, i, o);
      }
   }
   else if (op == BOP_MRES)
      blk_mres_dev<false>(a, imb, F, F + a.Ub, F + 2 * a.Ub, T, mu, SWG);
   else
      blk_mres_dev<true>(a, imb, F, F + a.Ub, F + 2 * a.Ub, T, mu, SWG);
}

// Dw_bnd (dirac/Dw_bnd.c:801-1011): for the face points x of block n the
// Weyl field w = -1/2 U chi of the hopping term into the exterior neighbour
// y, chi the upper components of (1 -+ gamma_mu) s(x): faces -mu (y = x - mu)
// use U(y,mu), faces +mu (y = x + mu) U(x,mu)^+.  Faces on the open time
// boundary give zero and zero s at their points.  One work-group per call.
struct FaceOfs
{
   int o[9];
};

__global__ __launch_bounds__(SWG) void k_dw_bnd(SArgs a, int n, int k, const int *ipp, FaceOfs fo, float2 *W)
{
   const int *imb = a.imb + (size_t)n * a.bvol;
   float2 *S = a.scr + ((size_t)n * NBF + k) * a.Ub;
   float2 *Wn = W + (size_t)n * fo.o[8] * 6;
   for (int ifc = 0; ifc < 8; ifc++)
   {
      const int mu = ifc >> 1;
      const bool up = ifc & 1;
      for (int j = fo.o[ifc] + threadIdx.x; j < fo.o[ifc + 1]; j += SWG)
      {
         const int x = ipp[j], gx = imb[x];
         float2 w[6];
         if (mu == 0 && a.masked && a.bnd[gx])
         {
#pragma unroll
            for (int c = 0; c < 6; c++) w[c] = make_float2(0.0f, 0.0f);
            zero_spinor(S, x);
         }
         else
         {
            float2 sp[12], h[6], u[9];
            load_spinor((const float2 *)S, x, sp);
            // chi (spin components 1, 2) per face, from Dw_bnd.c:825-1009
#pragma unroll
            for (int c = 0; c < 3; c++)
            {
               const float2 s1 = sp[c], s2 = sp[3 + c], s3 = sp[6 + c], s4 = sp[9 + c];
               float2 a1, a2;
               if (mu == 0)
               {
                  a1 = up ? csub(s1, s3) : cadd(s1, s3);
                  a2 = up ? csub(s2, s4) : cadd(s2, s4);
               }
               else if (mu == 1)
               {
                  a1 = up ? csubi(s1, s4) : caddi(s1, s4);
                  a2 = up ? csubi(s2, s3) : caddi(s2, s3);
               }
               else if (mu == 2)
               {
                  a1 = up ? csub(s1, s4) : cadd(s1, s4);
                  a2 = up ? cadd(s2, s3) : csub(s2, s3);
               }
               else
               {
                  a1 = up ? csubi(s1, s3) : caddi(s1, s3);
                  a2 = up ? caddi(s2, s4) : csubi(s2, s4);
               }
               h[c] = cscale(a1, -0.5f);
               h[3 + c] = cscale(a2, -0.5f);
            }
            const int gy = a.nbr[(size_t)(up ? 2 * mu : 2 * mu + 1) * a.V + gx] & 0x7fffffff;
            const float2 *lp = blk_link(a, gx, gy, mu, up);
#pragma unroll
            for (int q = 0; q < 9; q++) u[q] = lp[q * SB];
            if (up)
            {
               su3_mul<float2, true>(u, h, w);
               su3_mul<float2, true>(u, h + 3, w + 3);
            }
            else
            {
               su3_mul<float2, false>(u, h, w);
               su3_mul<float2, false>(u, h + 3, w + 3);
            }
         }
#pragma unroll
         for (int c = 0; c < 6; c++) Wn[(size_t)j * 6 + c] = w[c];
      }
      // the time faces zero s on the open boundary before the other faces
      // read it (the faces are processed in the order -0, +0, ..., +3)
      if (ifc < 2) __syncthreads();
   }
}

template <typename C> __global__ void k_blk_copy(C *glob, C *blk, const int *imb, int i0, int i1, int to_blk)
{
   int i = i0 + blockIdx.x * 256 + threadIdx.x;
   if (i >= i1) return;
   int gx = imb[i];
   C v[12];
   if (to_blk)
   {
      load_spinor((const C *)glob, gx, v);
      store_spinor(blk, i, v);
   }
   else
   {
      load_spinor((const C *)blk, i, v);
      store_spinor(glob, gx, v);
   }
}

// ---------------------------------------------------------------------------
// host side: block grid
// ---------------------------------------------------------------------------

static int lexb(const int *bs, const int *y) { return y[3] + bs[3] * (y[2] + bs[2] * (y[1] + bs[1] * y[0])); }

static void drop_sap_graph();

static void free_grid()
{
   drop_sap_graph();
   dev_free(G.d_imb); dev_free(G.d_bnb);
   for (int c = 0; c < 2; c++) { dev_free(G.d_face[c]); dev_free(G.d_fmask[c]); G.d_face[c] = nullptr; G.d_fmask[c] = nullptr; }
   dev_free(G.d_swb); dev_free(G.d_scr); dev_free(G.d_s0g); dev_free(G.d_bs);
   dev_free(G.d_u16); dev_free(G.d_swb16);
   for (int k = 0; k < 3; k++) { dev_free(G.d_h16[k]); dev_free(G.d_hsc[k]); dev_free(G.d_h64[k]); }
   dev_free(G.d_swb64); dev_free(G.d_scr64); dev_free(G.d_s0gm); dev_free(G.d_bsd);
   dev_free(G.d_ipp); dev_free(G.d_bw);
   G = SapGrid();
}

static bool build_grid(const char *where)
{
   Lib &L = lib();
   if (!require_geometry(where)) return false;
   if (L.sap_ncy == 0)
   {
      fail(where, "SAP parameters are not set");
      return false;
   }
   if (G.ready && memcmp(G.bs, L.sap_bs, sizeof(G.bs)) == 0) return true;
   if (G.ready) free_grid();
   const int *bs = L.sap_bs;
   int nbl[4], cb[4], nbs[4];
   for (int mu = 0; mu < 4; mu++) nbl[mu] = L.g.L[mu] / bs[mu];
   memcpy(G.bs, bs, sizeof(G.bs));
   G.nb = nbl[0] * nbl[1] * nbl[2] * nbl[3];
   G.nbh = G.nb / 2;
   G.bvol = bs[0] * bs[1] * bs[2] * bs[3];
   G.bh = G.bvol / 2;
   G.Ub = units_for(G.bvol, 12);

   // block-local point order (blk_geometry.c:56-128): cache blocks, even first
   std::vector<int> bipt(G.bvol);
   cb[0] = bs[0];
   for (int mu = 1; mu < 4; mu++) cb[mu] = (bs[mu] % 4 == 0) ? 4 : ((bs[mu] % 3 == 0) ? 3 : 2);
   for (int mu = 0; mu < 4; mu++) nbs[mu] = bs[mu] / cb[mu];
   {
      int ofs[2] = {0, G.bh}, j[4], x[4], y[4];
      for (j[0] = 0; j[0] < nbs[0]; j[0]++)
      for (j[1] = 0; j[1] < nbs[1]; j[1]++)
      for (j[2] = 0; j[2] < nbs[2]; j[2]++)
      for (j[3] = 0; j[3] < nbs[3]; j[3]++)
         for (x[0] = 0; x[0] < cb[0]; x[0]++)
         for (x[1] = 0; x[1] < cb[1]; x[1]++)
         for (x[2] = 0; x[2] < cb[2]; x[2]++)
         for (x[3] = 0; x[3] < cb[3]; x[3]++)
         {
            for (int mu = 0; mu < 4; mu++) y[mu] = j[mu] * cb[mu] + x[mu];
            bipt[lexb(bs, y)] = ofs[(y[0] + y[1] + y[2] + y[3]) & 1]++;
         }
   }
   std::vector<int> bnb(8 * (size_t)G.bvol), bcoord(4 * (size_t)G.bvol);
   {
      int y[4];
      for (y[0] = 0; y[0] < bs[0]; y[0]++)
      for (y[1] = 0; y[1] < bs[1]; y[1]++)
      for (y[2] = 0; y[2] < bs[2]; y[2]++)
      for (y[3] = 0; y[3] < bs[3]; y[3]++)
      {
         int i = bipt[lexb(bs, y)];
         for (int mu = 0; mu < 4; mu++) bcoord[4 * (size_t)i + mu] = y[mu];
         for (int mu = 0; mu < 4; mu++)
         {
            int z[4];
            memcpy(z, y, sizeof(z));
            z[mu] = y[mu] + 1;
            bnb[(size_t)(2 * mu) * G.bvol + i] = (z[mu] < bs[mu]) ? bipt[lexb(bs, z)] : G.bvol;
            z[mu] = y[mu] - 1;
            bnb[(size_t)(2 * mu + 1) * G.bvol + i] = (z[mu] >= 0) ? bipt[lexb(bs, z)] : G.bvol;
         }
      }
   }
   // blocks: colour-0 blocks first, lexicographic (blk_grid.c:63-113)
   G.imb.assign((size_t)G.nb * G.bvol, 0);
   G.bo.assign(4 * (size_t)G.nb, 0);
   std::vector<int> colour_of_site(L.g.V, 0), inface(L.g.V, 0);
   std::vector<unsigned char> fm(L.g.V, 0);
   int ib = 0, m[4];
   for (int col = 0; col < 2; col++)
      for (m[0] = 0; m[0] < nbl[0]; m[0]++)
      for (m[1] = 0; m[1] < nbl[1]; m[1]++)
      for (m[2] = 0; m[2] < nbl[2]; m[2]++)
      for (m[3] = 0; m[3] < nbl[3]; m[3]++)
      {
         if (((m[0] + m[1] + m[2] + m[3]) & 1) != col) continue;
         for (int mu = 0; mu < 4; mu++) G.bo[4 * (size_t)ib + mu] = m[mu] * bs[mu];
         for (int i = 0; i < G.bvol; i++)
         {
            int g[4];
            for (int mu = 0; mu < 4; mu++) g[mu] = m[mu] * bs[mu] + bcoord[4 * (size_t)i + mu];
            int gx = L.g.ipt[g[3] + L.g.L[3] * (g[2] + L.g.L[2] * (g[1] + L.g.L[1] * g[0]))];
            G.imb[(size_t)ib * G.bvol + i] = gx;
            colour_of_site[gx] = col;
            unsigned char mk = 0;
            for (int mu = 0; mu < 4; mu++)
            {
               if (bnb[(size_t)(2 * mu) * G.bvol + i] == G.bvol) mk |= (unsigned char)(1 << (2 * mu));
               if (bnb[(size_t)(2 * mu + 1) * G.bvol + i] == G.bvol) mk |= (unsigned char)(2 << (2 * mu));
            }
            fm[gx] = mk;
         }
         ib++;
      }
   std::vector<int> face[2];
   std::vector<unsigned char> fmask[2];
   for (int ix = 0; ix < L.g.V; ix++)
      if (fm[ix])
      {
         face[colour_of_site[ix]].push_back(ix);
         fmask[colour_of_site[ix]].push_back(fm[ix]);
      }
   G.d_imb = (int *)dev_alloc(sizeof(int) * G.imb.size());
   // face points of the block boundaries in the order of bnd_geometry
   // (lattice/blk_geometry.c:241-313): odd points first, then even, each
   // point entered on the faces -mu/+mu whose neighbour lies outside
   {
      std::vector<int> fl[8];
      for (int ix = 0; ix < G.bvol; ix++)
      {
         const int iy = ix < G.bh ? ix + G.bh : ix - G.bh;
         for (int mu = 0; mu < 4; mu++)
         {
            if (bnb[(size_t)(2 * mu) * G.bvol + iy] == G.bvol) fl[2 * mu + 1].push_back(iy);
            if (bnb[(size_t)(2 * mu + 1) * G.bvol + iy] == G.bvol) fl[2 * mu].push_back(iy);
         }
      }
      std::vector<int> ipp;
      for (int ifc = 0; ifc < 8; ifc++)
      {
         G.fofs[ifc] = (int)ipp.size();
         ipp.insert(ipp.end(), fl[ifc].begin(), fl[ifc].end());
      }
      G.fofs[8] = (int)ipp.size();
      G.d_ipp = (int *)dev_alloc(sizeof(int) * ipp.size());
      HIP_OK(hipMemcpy(G.d_ipp, ipp.data(), sizeof(int) * ipp.size(), hipMemcpyHostToDevice));
   }
   G.d_bnb = (int *)dev_alloc(sizeof(int) * bnb.size());
   HIP_OK(hipMemcpy(G.d_imb, G.imb.data(), sizeof(int) * G.imb.size(), hipMemcpyHostToDevice));
   HIP_OK(hipMemcpy(G.d_bnb, bnb.data(), sizeof(int) * bnb.size(), hipMemcpyHostToDevice));
   for (int c = 0; c < 2; c++)
   {
      G.nface[c] = (int)face[c].size();
      G.d_face[c] = (int *)dev_alloc(sizeof(int) * (face[c].size() + 1));
      G.d_fmask[c] = (unsigned char *)dev_alloc(fmask[c].size() + 1);
      if (G.nface[c])
      {
         HIP_OK(hipMemcpy(G.d_face[c], face[c].data(), sizeof(int) * face[c].size(), hipMemcpyHostToDevice));
         HIP_OK(hipMemcpy(G.d_fmask[c], fmask[c].data(), fmask[c].size(), hipMemcpyHostToDevice));
      }
   }
   G.d_swb = (float2 *)dev_alloc(units_for(L.g.V, 36) * sizeof(float2));
   G.d_scr = (float2 *)dev_alloc((size_t)G.nbh * 3 * G.Ub * sizeof(float2));
   G.d_s0g = (float2 *)dev_alloc(units_for(L.g.V, 12) * sizeof(float2));
   HIP_OK(hipMemset(G.d_s0g, 0, units_for(L.g.V, 12) * sizeof(float2)));
   G.swb_set = -1;
   G.ready = true;
   return true;
}

// FP32 SW field of the block solver: single-precision copy of swd with the
// points of the set inverted in double precision first (map_sw2blk.c:53-100)
static int assign_swb(int set)
{
   Lib &L = lib();
   ensure_swd(0, 0);  // as sap_gcr.c:181-197: the SW field must not be inverted
   int h = L.g.V / 2, ifail = 0;
   size_t nun = units_for(L.g.V, 36);
   double2 *tmp = (double2 *)tmp_buffer(nun * sizeof(double2));
   HIP_OK(hipMemcpyAsync(tmp, L.d_swd, nun * sizeof(double2), hipMemcpyDeviceToDevice, L.stream));
   if (set == ALL_PTS || set == EVEN_PTS) ifail |= invert_pauli(tmp, tmp, 0, h, 0.0);
   if (set == ALL_PTS || set == ODD_PTS) ifail |= invert_pauli(tmp, tmp, h, h, 0.0);
   convert_f64_to_f32(tmp, G.d_swb, nun, L.stream);
   G.swb_version++;
   if (L.sap_prec == 2 || G.need64)
   {
      if (!G.d_swb64) G.d_swb64 = (double2 *)dev_alloc(nun * sizeof(double2));
      if (G.d_swb64)
      {
         HIP_OK(hipMemcpyAsync(G.d_swb64, tmp, nun * sizeof(double2), hipMemcpyDeviceToDevice, L.stream));
         G.swb64_version = G.swb_version;
      }
   }
   HIP_OK(hipStreamSynchronize(L.stream));
   G.swb_set = set;
   G.swb_epoch = L.swd_epoch;
   return ifail;
}

static SArgs sargs(float mu, int nmr)
{
   Lib &L = lib();
   SArgs a;
   memset(&a, 0, sizeof(a));  // compared bytewise by the graph cache
   a.u = L.d_u;
   a.sw = G.d_swb;
   a.imb = G.d_imb;
   a.bnb = G.d_bnb;
   a.nbr = L.d_nbr;
   a.bnd = L.d_bnd;
   a.V = L.g.V;
   a.h = L.g.V / 2;
   a.bvol = G.bvol;
   a.bh = G.bh;
   a.masked = (L.bc != 3);
   a.eoflg = L.eoflg;
   a.nmr = nmr;
   a.mu = mu;
   a.muo = L.eoflg ? 0.0f : mu;
   a.Ub = G.Ub;
   a.psi = a.rho = nullptr;
   a.s0g = G.d_s0g;
   a.scr = G.d_scr;
   a.n0 = 0;
   a.psi16 = G.d_h16[0];
   a.rho16 = G.d_h16[1];
   a.s0g16 = G.d_h16[2];
   a.psc = G.d_hsc[0];
   a.rsc = G.d_hsc[1];
   a.s0sc = G.d_hsc[2];
   a.pz = 0;
   return a;
}

static bool half_variant()
{
   return OQH_SAP_REG && G.bvol == RBV && lib().sap_prec == 1 && G.d_h16[0] && G.d_h16[1] && G.d_h16[2] && G.d_u16 &&
          G.d_swb16;
}

static bool reg_kernel() { return OQH_SAP_REG && G.bvol == RBV; }

// block solves of the blocks of colour ic; the register kernel first adds
// the face terms from the other colour's corrections if pull is set
static void sap_blocks(const SArgs &a0, int ic, int isolv, int pull)
{
   Lib &L = lib();
   SArgs a = a0;
   a.n0 = ic ? G.nbh : 0;
   a.pull = pull;
   if (a.nsrc)
   {
      // several sources: the FP32 register kernel only
      const unsigned nb = (unsigned)(G.nbh * a.nsrc);
      if (isolv && a.masked)
         k_sap_reg<1, float2, true><<<nb, RBV, 0, L.stream>>>(a, nullptr, nullptr);
      else if (isolv)
         k_sap_reg<1, float2, false><<<nb, RBV, 0, L.stream>>>(a, nullptr, nullptr);
      else if (a.masked)
         k_sap_reg<0, float2, true><<<nb, RBV, 0, L.stream>>>(a, nullptr, nullptr);
      else
         k_sap_reg<0, float2, false><<<nb, RBV, 0, L.stream>>>(a, nullptr, nullptr);
      return;
   }
   if (half_variant())
   {
      if (isolv && a.masked)
         k_sap_reg<1, __half2, true><<<G.nbh, RBV, 0, L.stream>>>(a, G.d_u16, G.d_swb16);
      else if (isolv)
         k_sap_reg<1, __half2, false><<<G.nbh, RBV, 0, L.stream>>>(a, G.d_u16, G.d_swb16);
      else if (a.masked)
         k_sap_reg<0, __half2, true><<<G.nbh, RBV, 0, L.stream>>>(a, G.d_u16, G.d_swb16);
      else
         k_sap_reg<0, __half2, false><<<G.nbh, RBV, 0, L.stream>>>(a, G.d_u16, G.d_swb16);
   }
   else if (reg_kernel())
   {
      if (isolv && a.masked)
         k_sap_reg<1, float2, true><<<G.nbh, RBV, 0, L.stream>>>(a, nullptr, nullptr);
      else if (isolv)
         k_sap_reg<1, float2, false><<<G.nbh, RBV, 0, L.stream>>>(a, nullptr, nullptr);
      else if (a.masked)
         k_sap_reg<0, float2, true><<<G.nbh, RBV, 0, L.stream>>>(a, nullptr, nullptr);
      else
         k_sap_reg<0, float2, false><<<G.nbh, RBV, 0, L.stream>>>(a, nullptr, nullptr);
   }
   else if (G.bvol >= SWG_MAX)
   {
      // large blocks (6^4, 8^4): 1024 threads per block, the few blocks of a
      // small lattice would otherwise leave most of the chip idle
      if (isolv)
         k_sap_blocks<1, float2, SWG_MAX><<<G.nbh, SWG_MAX, 0, L.stream>>>(a);
      else
         k_sap_blocks<0, float2, SWG_MAX><<<G.nbh, SWG_MAX, 0, L.stream>>>(a);
   }
   else if (isolv)
      k_sap_blocks<1><<<G.nbh, SWG, 0, L.stream>>>(a);
   else
      k_sap_blocks<0><<<G.nbh, SWG, 0, L.stream>>>(a);
}

// face terms from the corrections of the colour-(1-c) blocks into rho on
// the faces of the colour-c blocks
static void sap_push_faces(const SArgs &a, int c)
{
   Lib &L = lib();
   if (!G.nface[c]) return;
   unsigned nbk = (G.nface[c] + 255) / 256;
   if (a.nsrc)
   {
      k_sap_faces<float2><<<dim3(nbk, a.nsrc), 256, 0, L.stream>>>(a, G.d_face[c], G.d_fmask[c], G.nface[c], nullptr);
      return;
   }
   if (half_variant())
      k_sap_faces<__half2><<<nbk, 256, 0, L.stream>>>(a, G.d_face[c], G.d_fmask[c], G.nface[c], G.d_u16);
   else
      k_sap_faces<float2><<<nbk, 256, 0, L.stream>>>(a, G.d_face[c], G.d_fmask[c], G.nface[c], nullptr);
}

// one Schwarz cycle (sap.c:720-790): colour 0, faces of colour 1, colour 1,
// faces of colour 0.  With the FP16 register kernel the face terms into a
// colour are added by its block kernel (pull), so the cycle needs no face
// kernel except the final one; first: rho is complete on entry (no pull for
// colour 0); finish: complete rho at the end (faces of colour 0).
static void sap_cycle(const SArgs &a, int isolv, bool first, bool finish)
{
   // the FP32 register kernel keeps the face kernel: pulling the face terms
   // (scattered loads at the start of every block solve) measured no faster
   // at two wavefronts per SIMD (2.54 vs 2.49 ms per cycle at 48^4)
   if (half_variant() && !a.nsrc)
   {
      sap_blocks(a, 0, isolv, !first);
      sap_blocks(a, 1, isolv, 1);
      if (finish) sap_push_faces(a, 0);
   }
   else
   {
      sap_blocks(a, 0, isolv, 0);
      sap_push_faces(a, 1);
      sap_blocks(a, 1, isolv, 0);
      sap_push_faces(a, 0);
   }
}

// FP32 field <-> FP16 field with per-site scales
__global__ void k_s32_to_h16(const float2 *__restrict__ f, __half2 *h, float *sc, int V)
{
   int x = blockIdx.x * 256 + threadIdx.x;
   if (x >= V) return;
   float2 v[12];
   load_spinor(f, x, v);
   store_h16(h, sc, x, v);
}

__global__ void k_h16_to_s32(const __half2 *h, const float *sc, float2 *__restrict__ f, int V)
{
   int x = blockIdx.x * 256 + threadIdx.x;
   if (x >= V) return;
   float2 v[12];
   load_h16(h, sc, x, v);
   store_spinor(f, x, v);
}

// ---------------------------------------------------------------------------
// FP64 variant: the generic block solver on double fields, FP64 links and
// the FP64 block SW field
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_sap_faces64(SArgsD a, const int *face, const unsigned char *fmask, int nf)
{
   int k = blockIdx.x * 256 + threadIdx.x;
   if (k >= nf) return;
   int y = face[k];
   int m = fmask[k];
   double2 acc[12], sp[12];
   zero12(acc);
   int odd = y >= a.h;
#define OQH_FACE64(MU)                                                                 \
   if (m & (1 << (2 * MU)))                                                           \
   {                                                                                  \
      int x = a.nbr[(size_t)(2 * MU) * a.V + y] & 0x7fffffff;                         \
      load_spinor((const double2 *)a.s0g, x, sp);                                     \
      const double2 *up = odd ? link_ptr(a.u, y - a.h, 2 * MU) : link_ptr(a.u, x - a.h, 2 * MU + 1); \
      hop_one<double2, MU, true>(sp, up, acc);                                        \
   }                                                                                  \
   if (m & (2 << (2 * MU)))                                                           \
   {                                                                                  \
      int x = a.nbr[(size_t)(2 * MU + 1) * a.V + y] & 0x7fffffff;                     \
      load_spinor((const double2 *)a.s0g, x, sp);                                     \
      const double2 *up = odd ? link_ptr(a.u, y - a.h, 2 * MU + 1) : link_ptr(a.u, x - a.h, 2 * MU); \
      hop_one<double2, MU, false>(sp, up, acc);                                       \
   }
   OQH_FACE64(0)
   OQH_FACE64(1)
   OQH_FACE64(2)
   OQH_FACE64(3)
#undef OQH_FACE64
   double2 r[12];
   load_spinor((const double2 *)a.rho, y, r);
#pragma unroll
   for (int c = 0; c < 12; c++) r[c] = cadd(r[c], cscale(acc[c], 0.5));
   store_spinor(a.rho, y, r);
}

__global__ void k_s32_to_s64(const float2 *__restrict__ a, double2 *__restrict__ b, size_t n)
{
   size_t k = (size_t)blockIdx.x * 256 + threadIdx.x;
   if (k < n) b[k] = make_double2(a[k].x, a[k].y);
}

// phi = psi, chi = rho_in - rho (FP64 fields rounded to FP32)
__global__ void k_mop64_out(const double2 *__restrict__ psi, const double2 *__restrict__ rho,
                            const float2 *__restrict__ rin, float2 *__restrict__ phi, float2 *__restrict__ chi, size_t n)
{
   size_t k = (size_t)blockIdx.x * 256 + threadIdx.x;
   if (k >= n) return;
   phi[k] = make_float2((float)psi[k].x, (float)psi[k].y);
   chi[k] = make_float2((float)((double)rin[k].x - rho[k].x), (float)((double)rin[k].y - rho[k].y));
}

static SArgsD sargs64(double mu, int nmr)
{
   Lib &L = lib();
   SArgsD a;
   memset(&a, 0, sizeof(a));
   a.u = L.d_ud;
   a.sw = G.d_swb64;
   a.imb = G.d_imb;
   a.bnb = G.d_bnb;
   a.nbr = L.d_nbr;
   a.bnd = L.d_bnd;
   a.V = L.g.V;
   a.h = L.g.V / 2;
   a.bvol = G.bvol;
   a.bh = G.bh;
   a.masked = (L.bc != 3);
   a.eoflg = L.eoflg;
   a.nmr = nmr;
   a.mu = mu;
   a.muo = L.eoflg ? 0.0 : mu;
   a.Ub = G.Ub;
   a.psi = G.d_h64[0];
   a.rho = G.d_h64[1];
   a.s0g = G.d_h64[2];
   a.scr = G.d_scr64;
   return a;
}

static bool f64_variant() { return lib().sap_prec == 2 && G.d_h64[0] && G.d_scr64 && G.d_swb64; }

// one Schwarz cycle on the FP64 fields psi64, rho64
static void sap_cycle64(const SArgsD &a0, int isolv)
{
   Lib &L = lib();
   SArgsD a = a0;
   for (int ic = 0; ic < 2; ic++)
   {
      a.n0 = ic ? G.nbh : 0;
      if (isolv)
         k_sap_blocks<1><<<G.nbh, SWG, 0, L.stream>>>(a);
      else
         k_sap_blocks<0><<<G.nbh, SWG, 0, L.stream>>>(a);
      int c = 1 - ic;
      if (G.nface[c])
         k_sap_faces64<<<(G.nface[c] + 255) / 256, 256, 0, L.stream>>>(a, G.d_face[c], G.d_fmask[c], G.nface[c]);
   }
}

// the SAP preconditioner of sap_gcr in the FP64 variant: ncy cycles on FP64
// fields from psi = 0, rho = rho_in; phi = psi, chi = rho_in - rho
bool sap_mop64(double mu, int isolv, int nmr, int ncy, const float2 *rho, float2 *phi, float2 *chi)
{
   if (!f64_variant()) return false;
   Lib &L = lib();
   const size_t nu = units_for(L.g.V, 12);
   const unsigned nbk = (unsigned)((nu + 255) / 256);
   SArgsD a = sargs64(mu, nmr);
   HIP_OK(hipMemsetAsync(G.d_h64[0], 0, nu * sizeof(double2), L.stream));
   k_s32_to_s64<<<nbk, 256, 0, L.stream>>>(rho, G.d_h64[1], nu);
   for (int n = 0; n < ncy; n++) sap_cycle64(a, isolv);
   k_mop64_out<<<nbk, 256, 0, L.stream>>>(G.d_h64[0], G.d_h64[1], rho, phi, chi, nu);
   return true;
}

// one SAP cycle on device fields (no synchronisation); the FP16 variant
// converts psi and rho to FP16 spinors and back
void sap_device(float mu, int isolv, int nmr, float2 *psi, float2 *rho)
{
   Lib &L = lib();
   if (f64_variant())
   {
      const size_t nu = units_for(L.g.V, 12);
      const unsigned nbk = (unsigned)((nu + 255) / 256);
      k_s32_to_s64<<<nbk, 256, 0, L.stream>>>(psi, G.d_h64[0], nu);
      k_s32_to_s64<<<nbk, 256, 0, L.stream>>>(rho, G.d_h64[1], nu);
      sap_cycle64(sargs64(mu, nmr), isolv);
      convert_f64_to_f32(G.d_h64[0], psi, nu, L.stream);
      convert_f64_to_f32(G.d_h64[1], rho, nu, L.stream);
      return;
   }
   SArgs a = sargs(mu, nmr);
   a.psi = psi;
   a.rho = rho;
   const int V = L.g.V;
   const unsigned nbk = (V + 255) / 256;
   if (half_variant())
   {
      k_s32_to_h16<<<nbk, 256, 0, L.stream>>>(psi, G.d_h16[0], G.d_hsc[0], V);
      k_s32_to_h16<<<nbk, 256, 0, L.stream>>>(rho, G.d_h16[1], G.d_hsc[1], V);
   }
   sap_cycle(a, isolv, true, true);
   if (half_variant())
   {
      k_h16_to_s32<<<nbk, 256, 0, L.stream>>>(G.d_h16[0], G.d_hsc[0], psi, V);
      k_h16_to_s32<<<nbk, 256, 0, L.stream>>>(G.d_h16[1], G.d_hsc[1], rho, V);
   }
}

// ncy Schwarz cycles (FP32) on n sources at once (n <= 16): psi[k] and
// rho[k] are full FP32 device fields; each launch takes all sources
bool sap_multi_device(float mu, int isolv, int nmr, int ncy, int n, float2 *const *psi, float2 *const *rho)
{
   Lib &L = lib();
   if (n < 1 || n > 16 || !reg_kernel()) return false;
   if (!G.d_s0gm)
   {
      G.d_s0gm = (float2 *)dev_alloc(16 * units_for(L.g.V, 12) * sizeof(float2));
      if (!G.d_s0gm) return false;
   }
   SArgs a = sargs(mu, nmr);
   a.nsrc = n;
   for (int k = 0; k < n; k++)
   {
      a.pl[k] = psi[k];
      a.rl[k] = rho[k];
      a.sl[k] = G.d_s0gm + (size_t)k * units_for(L.g.V, 12);
   }
   for (int c = 0; c < ncy; c++) sap_cycle(a, isolv, true, true);
   return true;
}

// the FP16 residual of the last Schwarz cycles as an FP32 field
void sap_mixed_residual(float2 *r)
{
   Lib &L = lib();
   const int V = L.g.V;
   k_h16_to_s32<<<(V + 255) / 256, 256, 0, L.stream>>>(G.d_h16[1], G.d_hsc[1], r, V);
}

#ifndef OQH_SAP_GRAPH
#define OQH_SAP_GRAPH 1
#endif

// the Schwarz cycles of the FP16 preconditioner as one HIP graph: its kernel
// arguments depend only on rho (fixed within fgcr), the SAP parameters and
// the grid, so the graph is made when the same arguments come a second time
// in a row and replayed for every further Krylov vector; calls with
// changing arguments (the fields of dfl_modes) run eagerly
struct SapCall
{
   SArgs a;
   const float2 *rho = nullptr;
   int isolv = -1, ncy = -1, want_res = -1;
   bool same(const SapCall &o) const
   {
      return rho && rho == o.rho && isolv == o.isolv && ncy == o.ncy && want_res == o.want_res &&
             memcmp(&a, &o.a, sizeof(a)) == 0;
   }
};
static struct
{
   hipGraphExec_t exec = nullptr;
   SapCall made;  // the arguments of exec
   SapCall last;  // the arguments of the last eager call
} SG;

static void drop_sap_graph()
{
   if (SG.exec)
   {
      // a launch of the graph may still be running on the library stream
      HIP_OK(hipStreamSynchronize(lib().stream));
      HIP_OK(hipGraphExecDestroy(SG.exec));
   }
   SG.exec = nullptr;
   SG.made = SapCall();
   SG.last = SapCall();
}

// rho -> FP16, then ncy cycles from psi = 0 (enqueued on the library stream);
// the final face terms only if the residual is wanted
static void enqueue_mixed_cycles(const SArgs &a0, int isolv, int ncy, const float2 *rho, bool want_res)
{
   Lib &L = lib();
   const int V = L.g.V;
   SArgs a = a0;
   k_s32_to_h16<<<(V + 255) / 256, 256, 0, L.stream>>>(rho, G.d_h16[1], G.d_hsc[1], V);
   for (int n = 0; n < ncy; n++)
   {
      // the blocks of both colours see psi = 0 in the first cycle
      a.pz = (n == 0);
      sap_cycle(a, isolv, n == 0, n == ncy - 1 && want_res);
   }
}

// the SAP preconditioner of sap_gcr in the FP16 variant: ncy cycles on FP16
// fields starting from psi = 0, rho = rho_in, then phi = psi (sap_gcr.c:99-117
// with the fields kept in FP16).  The residual carried in FP16 no longer
// equals rho_in - D psi to FP32 accuracy, so the caller computes chi = D phi
// with the FP32 operator instead of taking rho_in - rho.
bool sap_mixed_mop(float mu, int isolv, int nmr, int ncy, const float2 *rho, float2 *phi, bool want_res)
{
   if (!half_variant()) return false;
   Lib &L = lib();
   const int V = L.g.V;
   SArgs a = sargs(mu, nmr);
   SapCall c;
   memset(&c.a, 0, sizeof(c.a));  // padding included, for the byte comparison
   c.a = a;
   c.rho = rho;
   c.isolv = isolv;
   c.ncy = ncy;
   c.want_res = (int)want_res;
   if (OQH_SAP_GRAPH && L.sap_graph && SG.exec && SG.made.same(c))
      HIP_OK(hipGraphLaunch(SG.exec, L.stream));
   else if (OQH_SAP_GRAPH && L.sap_graph && SG.last.same(c))
   {
      drop_sap_graph();
      hipGraph_t g;
      HIP_OK(hipStreamBeginCapture(L.stream, hipStreamCaptureModeThreadLocal));
      enqueue_mixed_cycles(a, isolv, ncy, rho, want_res);
      HIP_OK(hipStreamEndCapture(L.stream, &g));
      HIP_OK(hipGraphInstantiate(&SG.exec, g, nullptr, nullptr, 0));
      HIP_OK(hipGraphDestroy(g));
      SG.made = c;
      HIP_OK(hipGraphLaunch(SG.exec, L.stream));
   }
   else
   {
      enqueue_mixed_cycles(a, isolv, ncy, rho, want_res);
      SG.last = c;
   }
   k_h16_to_s32<<<(V + 255) / 256, 256, 0, L.stream>>>(G.d_h16[0], G.d_hsc[0], phi, V);
   return true;
}

__global__ void k_f32_to_f16(const float2 *__restrict__ a, __half2 *__restrict__ b, size_t n)
{
   size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
   if (i < n) b[i] = __float22half2_rn(a[i]);
}

// FP16 copies of the links and of the block SW field (FP16 variant)
static void ensure_half_fields()
{
   Lib &L = lib();
   size_t nu = units_for(L.g.V / 2, 72), ns = units_for(L.g.V, 36);
   if (!G.d_u16) G.d_u16 = (__half2 *)dev_alloc(nu * sizeof(__half2));
   if (!G.d_swb16) G.d_swb16 = (__half2 *)dev_alloc(ns * sizeof(__half2));
   for (int k = 0; k < 3; k++)
   {
      if (!G.d_h16[k]) G.d_h16[k] = (__half2 *)dev_alloc(units_for(L.g.V, 12) * sizeof(__half2));
      if (!G.d_hsc[k]) G.d_hsc[k] = (float *)dev_alloc((size_t)L.g.V * sizeof(float));
      if (!G.d_h16[k] || !G.d_hsc[k]) return;
   }
   if (!G.d_u16 || !G.d_swb16) return;
   if (G.u16_epoch != L.u_epoch)
   {
      k_f32_to_f16<<<(unsigned)((nu + 255) / 256), 256, 0, L.stream>>>(L.d_u, G.d_u16, nu);
      G.u16_epoch = L.u_epoch;
   }
   if (G.swb16_epoch != G.swb_version)
   {
      k_f32_to_f16<<<(unsigned)((ns + 255) / 256), 256, 0, L.stream>>>(G.d_swb, G.d_swb16, ns);
      G.swb16_epoch = G.swb_version;
   }
}

// grid, single-precision links and block SW field for the solver; returns
// false if the odd-point inversion is unsafe (reported as ifail=-3 by
// sap_gcr, as an error by sap)
bool sap_prepare(int isolv, const char *where, bool quiet)
{
   Lib &L = lib();
   if (!build_grid(where)) return false;
   if (!L.ud_set)
   {
      fail(where, "gauge field is not set");
      return false;
   }
   ensure_u();
   int want = isolv ? ODD_PTS : NO_PTS;
   if (!L.swd_valid || L.swd_inv[0] || L.swd_inv[1]) ensure_swd(0, 0);
   if (G.swb_set != want || G.swb_epoch != L.swd_epoch || (L.sap_prec == 2 && G.swb64_version != G.swb_version))
   {
      if (assign_swb(want))
      {
         G.swb_set = -1;
         if (!quiet) fail(where, "The inversion of the SW term was not safe");
         return false;
      }
   }
   if (L.sap_prec == 1 && G.bvol == RBV) ensure_half_fields();
   if (L.sap_prec == 2)
   {
      for (int k = 0; k < 3; k++)
         if (!G.d_h64[k]) G.d_h64[k] = (double2 *)dev_alloc(units_for(L.g.V, 12) * sizeof(double2));
      if (!G.d_scr64) G.d_scr64 = (double2 *)dev_alloc((size_t)G.nbh * 3 * G.Ub * sizeof(double2));
      if (!G.d_h64[0] || !G.d_h64[1] || !G.d_h64[2] || !G.d_scr64 || !G.d_swb64) return false;
   }
   return true;
}

int sap_grid_nblocks() { return G.ready ? G.nb : 0; }

// the grid belongs to the geometry: released with it
void sap_release()
{
   drop_sap_graph();
   if (G.ready || G.d_bs || G.d_bsd) free_grid();
}

}  // namespace oqh

using namespace oqh;

extern "C" void sap(float mu, int isolv, int nmr, spinor *psi, spinor *eta)
{
   Lib &L = lib();
   if (!sap_prepare(isolv, "sap [sap.c]", false)) return;
   int V = L.g.V;
   Res p = resolve(psi, F32, V, 0, true);
   Res r = resolve(eta, F32, V, 1, true);
   if (!p.dev || !r.dev) return;
   if (p.off || r.off)
   {
      fail("sap [sap.c]", "the fields must be full fields");
      return;
   }
   sap_device(mu, isolv, nmr, (float2 *)p.dev, (float2 *)r.dev);
   finish(p, F32, V, true);
   finish(r, F32, V, true);
}

// ---------------------------------------------------------------------------
// block-size autotuner (no counterpart in the reference)
// ---------------------------------------------------------------------------

// reproducible test field: uniform values in [-1, 1) from a hash of the index
__global__ void k_hash_fill(float2 *f, size_t n, unsigned long long seed)
{
   size_t k = (size_t)blockIdx.x * 256 + threadIdx.x;
   if (k >= n) return;
   unsigned long long z = seed + 0x9E3779B97F4A7C15ull * (k + 1);
   z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
   z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
   z ^= z >> 31;
   const float a = (float)(z & 0xffffff) * (1.0f / 8388608.0f) - 1.0f;
   const float b = (float)((z >> 24) & 0xffffff) * (1.0f / 8388608.0f) - 1.0f;
   f[k] = make_float2(a, b);
}

__global__ void k_sq_partial(const float2 *f, size_t n, double *part)
{
   __shared__ double sh[256];
   double s = 0.0;
   for (size_t k = (size_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (size_t)gridDim.x * 256)
      s += (double)f[k].x * f[k].x + (double)f[k].y * f[k].y;
   sh[threadIdx.x] = s;
   __syncthreads();
   for (int m = 128; m > 0; m >>= 1)
   {
      if (threadIdx.x < m) sh[threadIdx.x] += sh[threadIdx.x + m];
      __syncthreads();
   }
   if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

static double field_norm(const float2 *f, size_t n, double *d_part, double *h_part)
{
   Lib &L = lib();
   k_sq_partial<<<256, 256, 0, L.stream>>>(f, n, d_part);
   HIP_OK(hipMemcpyAsync(h_part, d_part, 256 * sizeof(double), hipMemcpyDeviceToHost, L.stream));
   HIP_OK(hipStreamSynchronize(L.stream));
   double s = 0.0;
   for (int b = 0; b < 256; b++) s += h_part[b];
   return sqrt(s);
}

// legal block sizes (sap_parms.c / README.block: even, >= 4, dividing the
// lattice with an even number of blocks per axis), spatially isotropic, at
// most maxvol points
static std::vector<std::array<int, 4>> block_candidates(const int *Lx, int maxvol)
{
   std::vector<std::array<int, 4>> out;
   auto legal = [](int l, int b) { return b >= 4 && b % 2 == 0 && l % b == 0 && (l / b) % 2 == 0; };
   for (int b0 = 4; b0 <= Lx[0]; b0 += 2)
      for (int bs = 4; bs <= std::min(Lx[1], std::min(Lx[2], Lx[3])); bs += 2)
      {
         if (!legal(Lx[0], b0) || !legal(Lx[1], bs) || !legal(Lx[2], bs) || !legal(Lx[3], bs)) continue;
         if ((long)b0 * bs * bs * bs > maxvol) continue;
         out.push_back({b0, bs, bs, bs});
      }
   return out;
}

extern "C" void oqh_set_sap_graph(int on)
{
   if (on != 0 && on != 1)
   {
      fail("oqh_set_sap_graph", "the argument must be 0 or 1");
      return;
   }
   lib().sap_graph = on;
}

extern "C" void oqh_set_sap_prec(int prec)
{
   if (prec != 0 && prec != 1 && prec != 2)
   {
      fail("oqh_set_sap_prec", "the precision must be 0 (FP32), 1 (FP16 storage) or 2 (FP64)");
      return;
   }
   lib().sap_prec = prec;
}

extern "C" void alloc_bgr(blk_grid_t grid)
{
   if (grid != SAP_BLOCKS)
   {
      fail("alloc_bgr [blk_grid.c]", "only the SAP_BLOCKS grid is provided");
      return;
   }
   build_grid("alloc_bgr [blk_grid.c]");
}

extern "C" void assign_ud2ubgr(blk_grid_t grid)
{
   (void)grid;
   Lib &L = lib();
   if (!require_geometry("assign_ud2ubgr [map_u2blk.c]")) return;
   L.u_valid = false;
   ensure_u();
}

extern "C" int assign_swd2swbgr(blk_grid_t grid, ptset_t set)
{
   if (grid != SAP_BLOCKS || !build_grid("assign_swd2swbgr [map_sw2blk.c]")) return 1;
   return assign_swb((int)set);
}

static bool blk_check(int n, int k, const char *where)
{
   if (!build_grid(where)) return false;
   if (n < 0 || n >= G.nb || k < 0 || k > 2)
   {
      fail(where, "Block number or field index is out of range");
      return false;
   }
   if (!G.d_bs) G.d_bs = (float2 *)dev_alloc((size_t)G.nb * NBF * G.Ub * sizeof(float2));
   return true;
}

static void blk_copy(int n, ptset_t set, int k, spinor *s, int to_blk, const char *where)
{
   Lib &L = lib();
   if (!blk_check(n, k, where)) return;
   Res r = resolve(s, F32, L.g.V, 0, true);
   if (!r.dev) return;
   int i0 = (set == ODD_PTS) ? G.bh : 0, i1 = (set == EVEN_PTS) ? G.bh : G.bvol;
   if (set == NO_PTS) i1 = i0;
   float2 *blk = G.d_bs + ((size_t)n * NBF + k) * G.Ub;
   if (i1 > i0)
      k_blk_copy<<<(i1 - i0 + 255) / 256, 256, 0, L.stream>>>((float2 *)r.dev, blk, G.d_imb + (size_t)n * G.bvol,
                                                              i0, i1, to_blk);
   finish(r, F32, L.g.V, !to_blk);
}

extern "C" void assign_s2sblk(blk_grid_t grid, int n, ptset_t set, spinor *s, int k)
{
   (void)grid;
   blk_copy(n, set, k, s, 1, "assign_s2sblk [map_s2blk.c]");
}

extern "C" void assign_sblk2s(blk_grid_t grid, int n, ptset_t set, int k, spinor *s)
{
   (void)grid;
   blk_copy(n, set, k, s, 0, "assign_sblk2s [map_s2blk.c]");
}

static void blk_op(int n, int op, int k, int l, float mu, int nmr, const char *where)
{
   Lib &L = lib();
   if (!blk_check(n, k, where) || !blk_check(n, l, where)) return;
   if (!L.ud_set)
   {
      fail(where, "gauge field is not set");
      return;
   }
   ensure_u();
   // the block SW field follows the SW term of the current gauge field and
   // parameters: it is (re)made with the last assigned point set (or the one
   // of the SAP parameters) when it is missing or older than the SW field
   if (G.swb_set < 0 || !L.swd_valid || G.swb_epoch != L.swd_epoch)
   {
      int set = (G.swb_set >= 0) ? G.swb_set : (L.sap_isolv ? ODD_PTS : NO_PTS);
      if (assign_swb(set))
      {
         G.swb_set = -1;
         fail(where, "The inversion of the SW term was not safe");
         return;
      }
   }
   SArgs a = sargs(mu, nmr);
   a.scr = G.d_bs;
   k_blk_op<<<1, SWG, 0, L.stream>>>(a, n, op, k, l, mu);
   HIP_OK(hipStreamSynchronize(L.stream));
}

extern "C" void Dw_blk(blk_grid_t grid, int n, float mu, int k, int l)
{
   (void)grid;
   if (k == l)
   {
      fail("Dw_blk [Dw.c]", "Input and output fields must be different");
      return;
   }
   blk_op(n, BOP_DW, k, l, mu, 0, "Dw_blk [Dw.c]");
}

extern "C" void Dwhat_blk(blk_grid_t grid, int n, float mu, int k, int l)
{
   (void)grid;
   if (k == l)
   {
      fail("Dwhat_blk [Dw.c]", "Input and output fields must be different");
      return;
   }
   blk_op(n, BOP_DWHAT, k, l, mu, 0, "Dwhat_blk [Dw.c]");
}

extern "C" void Dwee_blk(blk_grid_t grid, int n, float mu, int k, int l)
{
   (void)grid;
   blk_op(n, BOP_DWEE, k, l, mu, 0, "Dwee_blk [Dw.c]");
}

extern "C" void Dwoo_blk(blk_grid_t grid, int n, float mu, int k, int l)
{
   (void)grid;
   blk_op(n, BOP_DWOO, k, l, mu, 0, "Dwoo_blk [Dw.c]");
}

extern "C" void Dwoe_blk(blk_grid_t grid, int n, int k, int l)
{
   (void)grid;
   blk_op(n, BOP_DWOE, k, l, 0.0f, 0, "Dwoe_blk [Dw.c]");
}

extern "C" void Dweo_blk(blk_grid_t grid, int n, int k, int l)
{
   (void)grid;
   blk_op(n, BOP_DWEO, k, l, 0.0f, 0, "Dweo_blk [Dw.c]");
}

// ---------------------------------------------------------------------------
// double-precision block fields and operators (block/map_s2blk.c:
// assign_sd2sdblk, assign_sdblk2sd; dirac/Dw_dble.c:1891-2290): the same
// block kernels on FP64 block fields with the FP64 links and the FP64 block
// SW field (inverted on the point set of the single-precision one)
// ---------------------------------------------------------------------------

static bool blk_check64(int n, int k, const char *where)
{
   if (!blk_check(n, k, where)) return false;
   if (!G.d_bsd) G.d_bsd = (double2 *)dev_alloc((size_t)G.nb * NBF * G.Ub * sizeof(double2));
   if (!G.d_bsd) return false;
   return true;
}

static void blk_copy64(int n, ptset_t set, int k, spinor_dble *sd, int to_blk, const char *where)
{
   Lib &L = lib();
   if (!blk_check64(n, k, where)) return;
   Res r = resolve(sd, F64, L.g.V, 0, true);
   if (!r.dev) return;
   int i0 = (set == ODD_PTS) ? G.bh : 0, i1 = (set == EVEN_PTS) ? G.bh : G.bvol;
   if (set == NO_PTS) i1 = i0;
   double2 *blk = G.d_bsd + ((size_t)n * NBF + k) * G.Ub;
   if (i1 > i0)
      k_blk_copy<<<(i1 - i0 + 255) / 256, 256, 0, L.stream>>>((double2 *)r.dev, blk, G.d_imb + (size_t)n * G.bvol,
                                                              i0, i1, to_blk);
   finish(r, F64, L.g.V, !to_blk);
}

extern "C" void assign_sd2sdblk(blk_grid_t grid, int n, ptset_t set, spinor_dble *sd, int k)
{
   (void)grid;
   blk_copy64(n, set, k, sd, 1, "assign_sd2sdblk [map_s2blk.c]");
}

extern "C" void assign_sdblk2sd(blk_grid_t grid, int n, ptset_t set, int k, spinor_dble *sd)
{
   (void)grid;
   blk_copy64(n, set, k, sd, 0, "assign_sdblk2sd [map_s2blk.c]");
}

static void blk_op64(int n, int op, int k, int l, double mu, const char *where)
{
   Lib &L = lib();
   if (!blk_check64(n, k, where) || !blk_check64(n, l, where)) return;
   if (!L.ud_set)
   {
      fail(where, "gauge field is not set");
      return;
   }
   ensure_u();
   G.need64 = true;
   if (G.swb_set < 0 || !L.swd_valid || G.swb_epoch != L.swd_epoch || !G.d_swb64 ||
       G.swb64_version != G.swb_version)
   {
      int set = (G.swb_set >= 0) ? G.swb_set : (L.sap_isolv ? ODD_PTS : NO_PTS);
      if (assign_swb(set) || !G.d_swb64)
      {
         G.swb_set = -1;
         fail(where, "The inversion of the SW term was not safe");
         return;
      }
   }
   SArgsD a = sargs64(mu, 0);
   a.scr = G.d_bsd;
   k_blk_op<<<1, SWG, 0, L.stream>>>(a, n, op, k, l, mu);
   HIP_OK(hipStreamSynchronize(L.stream));
}

extern "C" void Dw_blk_dble(blk_grid_t grid, int n, double mu, int k, int l)
{
   (void)grid;
   if (k == l)
   {
      fail("Dw_blk_dble [Dw_dble.c]", "Input and output fields must be different");
      return;
   }
   blk_op64(n, BOP_DW, k, l, mu, "Dw_blk_dble [Dw_dble.c]");
}

extern "C" void Dwhat_blk_dble(blk_grid_t grid, int n, double mu, int k, int l)
{
   (void)grid;
   if (k == l)
   {
      fail("Dwhat_blk_dble [Dw_dble.c]", "Input and output fields must be different");
      return;
   }
   blk_op64(n, BOP_DWHAT, k, l, mu, "Dwhat_blk_dble [Dw_dble.c]");
}

extern "C" void Dwee_blk_dble(blk_grid_t grid, int n, double mu, int k, int l)
{
   (void)grid;
   blk_op64(n, BOP_DWEE, k, l, mu, "Dwee_blk_dble [Dw_dble.c]");
}

extern "C" void Dwoo_blk_dble(blk_grid_t grid, int n, double mu, int k, int l)
{
   (void)grid;
   blk_op64(n, BOP_DWOO, k, l, mu, "Dwoo_blk_dble [Dw_dble.c]");
}

extern "C" void Dwoe_blk_dble(blk_grid_t grid, int n, int k, int l)
{
   (void)grid;
   blk_op64(n, BOP_DWOE, k, l, 0.0, "Dwoe_blk_dble [Dw_dble.c]");
}

extern "C" void Dweo_blk_dble(blk_grid_t grid, int n, int k, int l)
{
   (void)grid;
   blk_op64(n, BOP_DWEO, k, l, 0.0, "Dweo_blk_dble [Dw_dble.c]");
}

// FP64 block field k of block n as a host array in block order (test access)
extern "C" void oqh_export_sdblk(int n, int k, spinor_dble *host)
{
   Lib &L = lib();
   if (!blk_check64(n, k, "oqh_export_sdblk")) return;
   size_t bytes = (size_t)G.bvol * sizeof(spinor_dble);
   void *aos = tmp_buffer(bytes);
   blocks_to_aos_f64(G.d_bsd + ((size_t)n * NBF + k) * G.Ub, (double2 *)aos, G.bvol, 12, L.stream);
   HIP_OK(hipMemcpyAsync(host, aos, bytes, hipMemcpyDeviceToHost, L.stream));
   HIP_OK(hipStreamSynchronize(L.stream));
}

extern "C" void Dw_bnd(blk_grid_t grid, int n, int k, int l)
{
   Lib &L = lib();
   const char *where = "Dw_bnd [Dw_bnd.c]";
   (void)grid;
   if (!blk_check(n, k, where)) return;
   if (l != 0)
   {
      fail(where, "Parameters are out of range");
      return;
   }
   if (!L.ud_set)
   {
      fail(where, "gauge field is not set");
      return;
   }
   ensure_u();
   if (!G.d_bw)
   {
      G.d_bw = (float2 *)dev_alloc(sizeof(float2) * 6 * (size_t)G.nb * G.fofs[8]);
      if (!G.d_bw) return;
   }
   SArgs a = sargs(0.0f, 0);
   a.scr = G.d_bs;
   FaceOfs fo;
   memcpy(fo.o, G.fofs, sizeof(fo.o));
   k_dw_bnd<<<1, SWG, 0, L.stream>>>(a, n, k, G.d_ipp, fo, G.d_bw);
   HIP_OK(hipStreamSynchronize(L.stream));
}

extern "C" void oqh_export_bnd(int n, int ifc, int l, weyl *w)
{
   Lib &L = lib();
   const char *where = "oqh_export_bnd";
   if (!blk_check(n, 0, where)) return;
   if (ifc < 0 || ifc > 7 || l != 0 || !G.d_bw)
   {
      fail(where, "no such boundary field (Dw_bnd has not been called or parameters out of range)");
      return;
   }
   const size_t off = ((size_t)n * G.fofs[8] + G.fofs[ifc]) * 6;
   const size_t cnt = (size_t)(G.fofs[ifc + 1] - G.fofs[ifc]) * 6;
   HIP_OK(hipMemcpyAsync(w, G.d_bw + off, cnt * sizeof(float2), hipMemcpyDeviceToHost, L.stream));
   HIP_OK(hipStreamSynchronize(L.stream));
}

extern "C" void blk_mres(int n, float mu, int nmr)
{
   blk_op(n, BOP_MRES, 0, 2, mu, nmr, "blk_mres [blk_solv.c]");
}

extern "C" void blk_eo_mres(int n, float mu, int nmr)
{
   blk_op(n, BOP_EO_MRES, 0, 2, mu, nmr, "blk_eo_mres [blk_solv.c]");
}

extern "C" void oqh_export_sblk(int n, int k, spinor *host)
{
   Lib &L = lib();
   if (!blk_check(n, k, "oqh_export_sblk")) return;
   size_t bytes = (size_t)G.bvol * sizeof(spinor);
   void *aos = tmp_buffer(bytes);
   blocks_to_aos_f32(G.d_bs + ((size_t)n * NBF + k) * G.Ub, (float2 *)aos, G.bvol, 12, L.stream);
   HIP_OK(hipMemcpyAsync(host, aos, bytes, hipMemcpyDeviceToHost, L.stream));
   HIP_OK(hipStreamSynchronize(L.stream));
}

extern "C" void oqh_get_imb(int n, int *imb)
{
   if (!build_grid("oqh_get_imb")) return;
   memcpy(imb, G.imb.data() + (size_t)n * G.bvol, sizeof(int) * G.bvol);
}

// Chooses the SAP block size with the least time per e-fold of residual
// reduction: for every candidate, ncy Schwarz cycles with the given isolv and
// nmr (current precision variant) on a reproducible test residual, timed
// with HIP events after one warm-up cycle.  On success only the block size of
// the SAP parameters changes (isolv, nmr and ncy keep their values; if no SAP
// parameters were set, they become the given ones); on failure nothing
// changes.  Returns the best time per e-fold in seconds (0 on failure).
extern "C" double oqh_sap_autotune(int isolv, int nmr, int ncy, int maxvol, int *bs_out)
{
   const char *where = "oqh_sap_autotune";
   Lib &L = lib();
   if (!require_geometry(where)) return 0.0;
   if (!L.ud_set)
   {
      fail(where, "gauge field is not set");
      return 0.0;
   }
   if (maxvol <= 0) maxvol = 4096;
   auto cand = block_candidates(L.g.L, maxvol);
   if (cand.empty())
   {
      fail(where, "no legal block size");
      return 0.0;
   }
   const size_t nu = units_for(L.g.V, 12);
   float2 *psi = (float2 *)dev_alloc(nu * sizeof(float2));
   float2 *rho = (float2 *)dev_alloc(nu * sizeof(float2));
   double *d_part = (double *)dev_alloc(256 * sizeof(double));
   double *h_part = nullptr;
   HIP_OK(hipHostMalloc((void **)&h_part, 256 * sizeof(double), 0));
   hipEvent_t e0, e1;
   HIP_OK(hipEventCreate(&e0));
   HIP_OK(hipEventCreate(&e1));
   double best = 0.0;
   std::array<int, 4> bbest = cand[0];
   const int keep[4] = {L.sap_bs[0], L.sap_bs[1], L.sap_bs[2], L.sap_bs[3]};
   const int keep_isolv = L.sap_isolv, keep_nmr = L.sap_nmr, keep_ncy = L.sap_ncy;
   for (const auto &bs : cand)
   {
      memcpy(L.sap_bs, bs.data(), sizeof(L.sap_bs));
      L.sap_isolv = isolv;
      L.sap_nmr = nmr;
      L.sap_ncy = ncy;
      if (!psi || !rho || !sap_prepare(isolv, where, true)) continue;
      const unsigned nbk = (unsigned)((nu + 255) / 256);
      // warm-up cycle
      HIP_OK(hipMemsetAsync(psi, 0, nu * sizeof(float2), L.stream));
      k_hash_fill<<<nbk, 256, 0, L.stream>>>(rho, nu, 12345);
      sap_device((float)0.0f, isolv, nmr, psi, rho);
      HIP_OK(hipMemsetAsync(psi, 0, nu * sizeof(float2), L.stream));
      k_hash_fill<<<nbk, 256, 0, L.stream>>>(rho, nu, 12345);
      const double r0 = field_norm(rho, nu, d_part, h_part);
      HIP_OK(hipEventRecord(e0, L.stream));
      for (int n = 0; n < ncy; n++) sap_device(0.0f, isolv, nmr, psi, rho);
      HIP_OK(hipEventRecord(e1, L.stream));
      HIP_OK(hipEventSynchronize(e1));
      float ms = 0.0f;
      HIP_OK(hipEventElapsedTime(&ms, e0, e1));
      const double r1 = field_norm(rho, nu, d_part, h_part);
      if (!(r1 > 0.0) || r1 >= r0) continue;
      const double score = 1e-3 * ms / log(r0 / r1);
      if (best == 0.0 || score < best)
      {
         best = score;
         bbest = bs;
      }
   }
   HIP_OK(hipEventDestroy(e0));
   HIP_OK(hipEventDestroy(e1));
   HIP_OK(hipHostFree(h_part));
   dev_free(psi);
   dev_free(rho);
   dev_free(d_part);
   if (keep_ncy != 0 || best == 0.0)
   {
      // SAP parameters set before keep isolv, nmr and ncy
      L.sap_isolv = keep_isolv;
      L.sap_nmr = keep_nmr;
      L.sap_ncy = keep_ncy;
   }
   if (best == 0.0)
   {
      memcpy(L.sap_bs, keep, sizeof(L.sap_bs));
      fail(where, "no block size reduced the residual");
      return 0.0;
   }
   memcpy(L.sap_bs, bbest.data(), sizeof(L.sap_bs));
   for (int mu = 0; mu < 4; mu++) bs_out[mu] = bbest[mu];
   return best;
}

