// dfl.hip -- deflated SAP+GCR (local deflation of the low quark modes) on
// gfx950.
//
// Replaces flags/dfl_parms.c (set_dfl_parms, set_dfl_pro_parms,
// set_dfl_gen_parms), dfl/dfl_geometry.c (the DFL block grid),
// dfl/dfl_subspace.c (dfl_subspace, dfl_s2vd, dfl_vd2s,
// dfl_renormalize_modes), little/Aw_gen.c, Aw_ops.c, Aw_dble.c (the little
// Dirac operator and its even-odd preconditioned form), dfl/ltl_gcr.c with
// linsolv/fgcr4vd.c and gcr4v.c (the solver of the little Dirac equation),
// dfl/dfl_modes.c (generation of the subspace by inverse iteration) and
// dfl/dfl_sap_gcr.c (the deflated solver).
//
// Layout.  The DFL blocks partition the lattice, so the Ns orthonormal basis
// vectors v_{n,k} of all blocks n are kept as Ns global single-precision
// spinor fields B_k (B_k = v_{n,k} on block n), in the library's site-block
// layout: the projections and the prolongation then stream whole fields, and
// the Dirac kernels act on basis vectors directly.  Blocks are numbered even
// blocks first (block parity = parity of the block coordinates), each colour
// lexicographically; a little vector holds Ns complex numbers per block in
// that order.  The little operator is kept per block as Ablk = (v_{n,k},
// (Dw + i mu gamma5 I_eo) v_{n,l}) and Ahop[ifc] = (v_{n,k}, Dw v_{m,l}) for
// the 8 neighbouring blocks m = inn[n][ifc] (ifc = 2 mu: -mu, 2 mu + 1: +mu),
// computed in double precision from the FP32 basis with FP64 links and SW
// term.  The even-odd form stores Ablk^-1 and Ablk^-1 Ahop[ifc], transposed
// ([k] fastest) so that the 64 lanes of a matrix-vector product read rows
// as whole lines; an FP32 copy serves the inner single-precision solver.
//
// The little solver is FGCR in double precision preconditioned by
// single-pass GCR in single precision, both deflated by the little modes, as
// ltl_gcr.c with little/ltl_modes.c.  Difference to the reference: the random
// start fields of the subspace generation come from the double-precision
// RANLUX streams (uniform in [-1/2, 1/2) like ranlxs minus 1/2) rather than
// from the single-precision generator.
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "dirac_dev.h"
#include "oqh_internal.h"

namespace oqh {
bool sap_prepare(int isolv, const char *where, bool quiet);
void sap_device(float mu, int isolv, int nmr, float2 *psi, float2 *rho);
bool sap_mixed_mop(float mu, int isolv, int nmr, int ncy, const float2 *rho, float2 *phi, bool want_res);
}

using namespace oqh;

namespace {

constexpr int DWG = 256;     // work-group of the block kernels
constexpr int NS_MAX = 40;   // largest Ns (one lane per basis vector; Gauss-Jordan of Ns x 2Ns in LDS)

struct Dfl
{
   // parameters (flags/dfl_parms.c)
   int bs[4] = {0, 0, 0, 0}, Ns = 0;
   int nkv = 0, nmx = 0, nmx_gcr = 0;
   double res = 0.0, res_gcr = 0.0;
   int ninv = 0, nmr = 0, ncy = 0;
   double kappa = 0.0, m0 = 0.0, mu = 0.0;
   // grid
   bool grid = false;
   int nb = 0, nbh = 0, bvol = 0;
   int *d_bpts = nullptr;    // [nb][bvol] points of each block
   int *d_inn = nullptr;     // [nb][8] neighbouring blocks
   int *d_blk = nullptr;     // [V] block of every point
   std::vector<int> inn;     // host copy
   std::vector<int> bpts;    // host copy
   // subspace
   float2 *d_basis = nullptr;  // Ns fields of units_for(V,12)
   size_t fu = 0;              // units per field
   bool basis_ok = false;
   float2 *d_vmds = nullptr;   // [Ns][nb][Ns] coordinates of the generating modes
   double2 *d_s64 = nullptr;   // FP64 work field of the block Gram-Schmidt
   // little operator
   double2 *d_Ablk = nullptr;  // [nb][Ns][Ns]  (row k, column l)
   double2 *d_Ahop = nullptr;  // [nb][8][Ns][Ns]
   double2 *d_AinvT = nullptr; // [nb][Ns][Ns]  Ablk^-1 transposed ([l][k])
   double2 *d_MT = nullptr;    // [nb][8][Ns][Ns] (Ablk^-1 Ahop)^T
   float2 *d_MT32 = nullptr;   // its FP32 copy
   int *d_unsafe = nullptr;
   bool aw_ok = false, awhat_ok = false;
   double aw_mu = 0.0;
   int aw_ud = -1, aw_swd = -1, aw_eo = -1;
   // little vectors (device): work space of the little solver
   std::vector<double2 *> wvd;
   std::vector<float2 *> wv;
   double *d_red = nullptr, *h_red = nullptr;  // reductions
   // little modes (little/ltl_modes.c): Ns orthonormal vectors w_l on the
   // even blocks, Ahat w_l, both in FP64 and FP32, and M = (w, Ahat w)^-1
   bool lm_ok = false;
   size_t ldv = 0;
   double2 *d_lw = nullptr, *d_law = nullptr;
   float2 *d_lw32 = nullptr, *d_law32 = nullptr;
   double2 *d_lM = nullptr;     // [Ns][Ns] row-major
   double2 *d_lc = nullptr;     // [Ns] products, then [2][Ns] coefficients d, -d
   double2 *d_ld = nullptr;
   double2 *d_lB = nullptr;     // [Ns][Ns] (w_l, Ahat w_j) at column j
   double2 *d_lpart = nullptr;  // [Ns][LNBK] work-group shares
   double *d_lbc = nullptr;     // [Ns][4] Gram-Schmidt norms
   // Gram-Schmidt of the global modes (dfl_renormalize_modes)
   float2 *d_gcoef = nullptr, *d_ggram = nullptr, *d_gscr = nullptr;
   int gcap = 0;
};

Dfl D;

void dfl_free_grid();


// ---------------------------------------------------------------------------
// the DFL block grid (dfl/dfl_geometry.c, block/blk_grid.c)
// ---------------------------------------------------------------------------

bool dfl_grid(const char *where)
{
   Lib &L = lib();
   if (!require_geometry(where)) return false;
   if (D.Ns == 0)
   {
      fail(where, "Deflation subspace parameters are not set");
      return false;
   }
   if (D.grid) return true;
   const int *bs = D.bs;
   int nbl[4];
   for (int mu = 0; mu < 4; mu++) nbl[mu] = L.g.L[mu] / bs[mu];
   D.nb = nbl[0] * nbl[1] * nbl[2] * nbl[3];
   D.nbh = D.nb / 2;
   D.bvol = bs[0] * bs[1] * bs[2] * bs[3];
   // block numbers: even blocks first, lexicographic within a colour
   std::vector<int> bnum(D.nb);
   int ib = 0, m[4];
   for (int col = 0; col < 2; col++)
      for (m[0] = 0; m[0] < nbl[0]; m[0]++)
      for (m[1] = 0; m[1] < nbl[1]; m[1]++)
      for (m[2] = 0; m[2] < nbl[2]; m[2]++)
      for (m[3] = 0; m[3] < nbl[3]; m[3]++)
      {
         if (((m[0] + m[1] + m[2] + m[3]) & 1) != col) continue;
         bnum[m[3] + nbl[3] * (m[2] + nbl[2] * (m[1] + nbl[1] * m[0]))] = ib++;
      }
   D.inn.assign(8 * (size_t)D.nb, 0);
   D.bpts.assign((size_t)D.nb * D.bvol, 0);
   std::vector<int> fill(D.nb, 0);
   for (m[0] = 0; m[0] < nbl[0]; m[0]++)
   for (m[1] = 0; m[1] < nbl[1]; m[1]++)
   for (m[2] = 0; m[2] < nbl[2]; m[2]++)
   for (m[3] = 0; m[3] < nbl[3]; m[3]++)
   {
      const int n = bnum[m[3] + nbl[3] * (m[2] + nbl[2] * (m[1] + nbl[1] * m[0]))];
      for (int mu = 0; mu < 4; mu++)
      {
         int q[4] = {m[0], m[1], m[2], m[3]};
         q[mu] = (m[mu] + nbl[mu] - 1) % nbl[mu];
         D.inn[8 * (size_t)n + 2 * mu] = bnum[q[3] + nbl[3] * (q[2] + nbl[2] * (q[1] + nbl[1] * q[0]))];
         q[mu] = (m[mu] + 1) % nbl[mu];
         D.inn[8 * (size_t)n + 2 * mu + 1] = bnum[q[3] + nbl[3] * (q[2] + nbl[2] * (q[1] + nbl[1] * q[0]))];
      }
   }
   int y[4];
   for (y[0] = 0; y[0] < L.g.L[0]; y[0]++)
   for (y[1] = 0; y[1] < L.g.L[1]; y[1]++)
   for (y[2] = 0; y[2] < L.g.L[2]; y[2]++)
   for (y[3] = 0; y[3] < L.g.L[3]; y[3]++)
   {
      int q[4];
      for (int mu = 0; mu < 4; mu++) q[mu] = y[mu] / bs[mu];
      const int n = bnum[q[3] + nbl[3] * (q[2] + nbl[2] * (q[1] + nbl[1] * q[0]))];
      const int ix = L.g.ipt[y[3] + L.g.L[3] * (y[2] + L.g.L[2] * (y[1] + L.g.L[1] * y[0]))];
      D.bpts[(size_t)n * D.bvol + fill[n]++] = ix;
   }
   D.d_bpts = (int *)dev_alloc(sizeof(int) * D.bpts.size());
   D.d_inn = (int *)dev_alloc(sizeof(int) * D.inn.size());
   D.d_blk = (int *)dev_alloc(sizeof(int) * L.g.V);
   D.fu = units_for(L.g.V, 12);
   D.d_basis = (float2 *)dev_alloc(sizeof(float2) * D.fu * D.Ns);
   D.d_vmds = (float2 *)dev_alloc(sizeof(float2) * (size_t)D.Ns * D.nb * D.Ns);
   D.d_s64 = (double2 *)dev_alloc(sizeof(double2) * D.fu);
   const size_t nm = (size_t)D.Ns * D.Ns;
   D.d_Ablk = (double2 *)dev_alloc(sizeof(double2) * nm * D.nb);
   D.d_Ahop = (double2 *)dev_alloc(sizeof(double2) * nm * 8 * D.nb);
   D.d_AinvT = (double2 *)dev_alloc(sizeof(double2) * nm * D.nb);
   D.d_MT = (double2 *)dev_alloc(sizeof(double2) * nm * 8 * D.nb);
   D.d_MT32 = (float2 *)dev_alloc(sizeof(float2) * nm * 8 * D.nb);
   D.d_unsafe = (int *)dev_alloc(sizeof(int));
   D.d_red = (double *)dev_alloc(sizeof(double) * 4 * 1024);
   HIP_OK(hipHostMalloc((void **)&D.h_red, sizeof(double) * 4 * 1024, 0));
   if (!D.d_bpts || !D.d_inn || !D.d_blk || !D.d_basis || !D.d_vmds || !D.d_s64 || !D.d_Ablk || !D.d_Ahop || !D.d_AinvT ||
       !D.d_MT || !D.d_MT32 || !D.d_unsafe || !D.d_red)
   {
      dfl_free_grid();
      return false;
   }
   HIP_OK(hipMemcpy(D.d_bpts, D.bpts.data(), sizeof(int) * D.bpts.size(), hipMemcpyHostToDevice));
   HIP_OK(hipMemcpy(D.d_inn, D.inn.data(), sizeof(int) * D.inn.size(), hipMemcpyHostToDevice));
   {
      std::vector<int> blk(L.g.V);
      for (int n = 0; n < D.nb; n++)
         for (int i = 0; i < D.bvol; i++) blk[D.bpts[(size_t)n * D.bvol + i]] = n;
      HIP_OK(hipMemcpy(D.d_blk, blk.data(), sizeof(int) * L.g.V, hipMemcpyHostToDevice));
   }
   HIP_OK(hipMemset(D.d_basis, 0, sizeof(float2) * D.fu * D.Ns));
   D.grid = true;
   return true;
}

// ---------------------------------------------------------------------------
// block reductions
// ---------------------------------------------------------------------------

// sum of a double2 over the work-group (DWG lanes), result in every lane
__device__ inline double2 wg_sum(double2 v, double2 *sh)
{
#pragma unroll
   for (int m = 32; m >= 1; m >>= 1)
   {
      v.x += __shfl_xor(v.x, m);
      v.y += __shfl_xor(v.y, m);
   }
   const int w = threadIdx.x >> 6;
   __syncthreads();
   if ((threadIdx.x & 63) == 0) sh[w] = v;
   __syncthreads();
   double2 s = make_double2(0.0, 0.0);
#pragma unroll
   for (int k = 0; k < DWG / 64; k++)
   {
      s.x += sh[k].x;
      s.y += sh[k].y;
   }
   return s;
}

// (a, b) of two spinors, a FP32 promoted, b of type C
// sum over the 64 lanes of a wave
__device__ inline double2 wave_red2(double2 d)
{
#pragma unroll
   for (int m = 32; m >= 1; m >>= 1)
   {
      d.x += __shfl_xor(d.x, m);
      d.y += __shfl_xor(d.y, m);
   }
   return d;
}

template <typename C> __device__ inline double2 sdot(const float2 *a, const C *b)
{
   double2 z = make_double2(0.0, 0.0);
#pragma unroll
   for (int c = 0; c < 12; c++)
   {
      const double ax = a[c].x, ay = a[c].y, bx = b[c].x, by = b[c].y;
      z.x += ax * bx + ay * by;
      z.y += ax * by - ay * bx;
   }
   return z;
}

// w[n][k] = (B_k, F) on block n (dfl_s2vd: the components of F along the
// deflation subspace); one work-group per block
template <typename C>
__global__ __launch_bounds__(DWG) void k_blk_proj(const float2 *basis, size_t fu, int Ns, const int *bpts, int bvol,
                                                   const C *F, double2 *w)
{
   __shared__ double2 sh[DWG / 64];
   const int n = blockIdx.x;
   const int *pts = bpts + (size_t)n * bvol;
   for (int k = 0; k < Ns; k++)
   {
      double2 z = make_double2(0.0, 0.0);
      for (int i = threadIdx.x; i < bvol; i += DWG)
      {
         const int x = pts[i];
         float2 v[12];
         C f[12];
         load_spinor(basis + (size_t)k * fu, x, v);
         load_spinor(F, x, f);
         const double2 d = sdot(v, f);
         z.x += d.x;
         z.y += d.y;
      }
      z = wg_sum(z, sh);
      if (threadIdx.x == 0) w[(size_t)n * Ns + k] = z;
   }
}

// the same for blocks of at most DWG points: the field is read once into
// registers and the Ns products are reduced per wave, then across the waves
template <typename C>
__global__ __launch_bounds__(DWG) void k_blk_proj_reg(const float2 *basis, size_t fu, int Ns, const int *bpts,
                                                       int bvol, const C *F, double2 *w)
{
   __shared__ double2 sh[DWG / 64][NS_MAX];
   const int n = blockIdx.x, i = threadIdx.x, wv = threadIdx.x >> 6;
   const bool act = i < bvol;
   const int x = act ? bpts[(size_t)n * bvol + i] : 0;
   C f[12];
   if (act) load_spinor(F, x, f);
   for (int k = 0; k < Ns; k++)
   {
      double2 d = make_double2(0.0, 0.0);
      if (act)
      {
         float2 v[12];
         load_spinor(basis + (size_t)k * fu, x, v);
         d = sdot(v, f);
      }
      d = wave_red2(d);
      if ((threadIdx.x & 63) == 0) sh[wv][k] = d;
   }
   __syncthreads();
   for (int k = threadIdx.x; k < Ns; k += DWG)
   {
      double2 z = make_double2(0.0, 0.0);
      for (int q = 0; q < DWG / 64; q++)
      {
         z.x += sh[q][k].x;
         z.y += sh[q][k].y;
      }
      w[(size_t)n * Ns + k] = z;
   }
}

// s(x) = sum_k (float) w[n(x)][k] B_k(x) (dfl_vd2s: single-precision
// arithmetic as mulc_spinor_add)
__global__ __launch_bounds__(DWG) void k_blk_prolong(const float2 *basis, size_t fu, int Ns, const int *bpts,
                                                      int bvol, const double2 *w, float2 *s)
{
   const int n = blockIdx.x;
   const int *pts = bpts + (size_t)n * bvol;
   for (int i = threadIdx.x; i < bvol; i += DWG)
   {
      const int x = pts[i];
      float2 acc[12];
#pragma unroll
      for (int c = 0; c < 12; c++) acc[c] = make_float2(0.0f, 0.0f);
      for (int k = 0; k < Ns; k++)
      {
         const double2 wd = w[(size_t)n * Ns + k];
         const float2 z = make_float2((float)wd.x, (float)wd.y);
         float2 v[12];
         load_spinor(basis + (size_t)k * fu, x, v);
#pragma unroll
         for (int c = 0; c < 12; c++)
            acc[c] = make_float2(acc[c].x + (z.x * v[c].x - z.y * v[c].y), acc[c].y + (z.x * v[c].y + z.y * v[c].x));
      }
      store_spinor(s, x, acc);
   }
}

// block Gram-Schmidt of mode l for blocks of at most DWG points: one point
// per lane, the FP64 work spinor in registers, so that only the basis vectors
// are read (twice per previous vector, the second time from the cache)
__global__ __launch_bounds__(DWG) void k_blk_gs_reg(float2 *basis, size_t fu, int Ns, const int *bpts, int bvol,
                                                     const float2 *mds, int l, float2 *vmds, int *degen)
{
   __shared__ double2 sh[DWG / 64];
   const int n = blockIdx.x;
   const int *pts = bpts + (size_t)n * bvol;
   float2 *vm = vmds + ((size_t)l * gridDim.x + n) * Ns;
   const int i = threadIdx.x;
   const bool act = i < bvol;
   const int x = act ? pts[i] : 0;
   double2 d[12];
   {
      float2 v[12];
      if (act) load_spinor(mds, x, v);
#pragma unroll
      for (int c = 0; c < 12; c++) d[c] = act ? make_double2(v[c].x, v[c].y) : make_double2(0.0, 0.0);
   }
   for (int j = 0; j < l; j++)
   {
      const float2 *Bj = basis + (size_t)j * fu;
      float2 v[12];
      double2 z = make_double2(0.0, 0.0);
      if (act)
      {
         load_spinor(Bj, x, v);
         z = sdot(v, d);
      }
      z = wg_sum(z, sh);
      if (threadIdx.x == 0) vm[j] = make_float2((float)z.x, (float)z.y);
      if (act)
      {
#pragma unroll
         for (int c = 0; c < 12; c++)
            d[c] = make_double2(d[c].x - (z.x * v[c].x - z.y * v[c].y), d[c].y - (z.x * v[c].y + z.y * v[c].x));
      }
   }
   double2 r2 = make_double2(0.0, 0.0);
#pragma unroll
   for (int c = 0; c < 12; c++) r2.x += d[c].x * d[c].x + d[c].y * d[c].y;
   r2 = wg_sum(r2, sh);
   const double r = sqrt(r2.x);
   if (threadIdx.x == 0)
   {
      vm[l] = make_float2((float)r, 0.0f);
      for (int j = l + 1; j < Ns; j++) vm[j] = make_float2(0.0f, 0.0f);
      if ((float)r == 0.0f) atomicOr(degen, 1);
   }
   const double f = r > 0.0 ? 1.0 / r : 0.0;
   if (act)
   {
      float2 v[12];
#pragma unroll
      for (int c = 0; c < 12; c++) v[c] = make_float2((float)(d[c].x * f), (float)(d[c].y * f));
      store_spinor(basis + (size_t)l * fu, x, v);
   }
}

// block Gram-Schmidt of mode l (dfl_subspace): S = mds_l in double, S -=
// (B_j, S) B_j for j < l (modified Gram-Schmidt), B_l = S / |S|; the
// coordinates of mds_l in the basis go to vm[n][0..Ns) (complex float)
__global__ __launch_bounds__(DWG) void k_blk_gs(float2 *basis, size_t fu, int Ns, const int *bpts, int bvol,
                                                 const float2 *mds, double2 *S, int l, float2 *vmds, int *degen)
{
   __shared__ double2 sh[DWG / 64];
   const int n = blockIdx.x;
   const int *pts = bpts + (size_t)n * bvol;
   float2 *vm = vmds + ((size_t)l * gridDim.x + n) * Ns;
   for (int i = threadIdx.x; i < bvol; i += DWG)
   {
      const int x = pts[i];
      float2 v[12];
      double2 d[12];
      load_spinor(mds, x, v);
#pragma unroll
      for (int c = 0; c < 12; c++) d[c] = make_double2(v[c].x, v[c].y);
      store_spinor(S, x, d);
   }
   __syncthreads();
   for (int j = 0; j < l; j++)
   {
      const float2 *Bj = basis + (size_t)j * fu;
      double2 z = make_double2(0.0, 0.0);
      for (int i = threadIdx.x; i < bvol; i += DWG)
      {
         const int x = pts[i];
         float2 v[12];
         double2 d[12];
         load_spinor(Bj, x, v);
         load_spinor((const double2 *)S, x, d);
         const double2 p = sdot(v, d);
         z.x += p.x;
         z.y += p.y;
      }
      z = wg_sum(z, sh);
      if (threadIdx.x == 0) vm[j] = make_float2((float)z.x, (float)z.y);
      for (int i = threadIdx.x; i < bvol; i += DWG)
      {
         const int x = pts[i];
         float2 v[12];
         double2 d[12];
         load_spinor(Bj, x, v);
         load_spinor((const double2 *)S, x, d);
#pragma unroll
         for (int c = 0; c < 12; c++)
            d[c] = make_double2(d[c].x - (z.x * v[c].x - z.y * v[c].y), d[c].y - (z.x * v[c].y + z.y * v[c].x));
         store_spinor(S, x, d);
      }
      __syncthreads();
   }
   double2 r2 = make_double2(0.0, 0.0);
   for (int i = threadIdx.x; i < bvol; i += DWG)
   {
      double2 d[12];
      load_spinor((const double2 *)S, pts[i], d);
#pragma unroll
      for (int c = 0; c < 12; c++) r2.x += d[c].x * d[c].x + d[c].y * d[c].y;
   }
   r2 = wg_sum(r2, sh);
   const double r = sqrt(r2.x);
   if (threadIdx.x == 0)
   {
      vm[l] = make_float2((float)r, 0.0f);
      for (int j = l + 1; j < Ns; j++) vm[j] = make_float2(0.0f, 0.0f);
      if ((float)r == 0.0f) atomicOr(degen, 1);
   }
   const double f = r > 0.0 ? 1.0 / r : 0.0;
   for (int i = threadIdx.x; i < bvol; i += DWG)
   {
      const int x = pts[i];
      double2 d[12];
      float2 v[12];
      load_spinor((const double2 *)S, x, d);
#pragma unroll
      for (int c = 0; c < 12; c++) v[c] = make_float2((float)(d[c].x * f), (float)(d[c].y * f));
      store_spinor(basis + (size_t)l * fu, x, v);
   }
}

// ---------------------------------------------------------------------------
// the little Dirac operator (little/Aw_gen.c, Aw_ops.c)
// ---------------------------------------------------------------------------

struct AwArgs
{
   const float2 *basis;
   size_t fu;
   int Ns, bvol, V, h, masked, eoflg;
   const int *bpts, *nbr, *blk_of_unused;
   const unsigned char *bnd;
   const double2 *ud, *swd;
   double mu;
   double2 *Ablk, *Ahop;
};

// spinor of B_l at y promoted to double (zero at open-bc boundary points)
__device__ inline void ld_basis64(const AwArgs &a, int l, int y, double2 *s)
{
   float2 v[12];
   load_spinor(a.basis + (size_t)l * a.fu, y, v);
   const bool z = a.masked && a.bnd[y];
#pragma unroll
   for (int c = 0; c < 12; c++) s[c] = z ? make_double2(0.0, 0.0) : make_double2(v[c].x, v[c].y);
}

// column l of Ablk and Ahop[ifc] for every block: for each block point x,
// Y(x) = ((Dw + i mu gamma5 I_eo) restricted to the block) B_l at x (q = 0)
// and Z_ifc(x) = the hopping term into x from the neighbouring block in
// direction ifc (q = 1 + ifc); Ablk[n][k][l] = (B_k, Y)_n,
// Ahop[n][ifc][k][l] = (B_k, Z_ifc)_n, all in double precision

// hopping term of direction MU (FWD: from x + mu, else from x - mu) of B_l
// into x, added with factor -1/2 to acc
template <int MU, bool FWD>
__device__ inline void aw_hop(const AwArgs &a, int l, int x, int y, double2 *acc)
{
   const double2 *up;
   if (FWD)
      up = (x >= a.h) ? link_ptr(a.ud, x - a.h, 2 * MU) : link_ptr(a.ud, y - a.h, 2 * MU + 1);
   else
      up = (x >= a.h) ? link_ptr(a.ud, x - a.h, 2 * MU + 1) : link_ptr(a.ud, y - a.h, 2 * MU);
   double2 sp[12], t[12];
   ld_basis64(a, l, y, sp);
#pragma unroll
   for (int c = 0; c < 12; c++) t[c] = make_double2(0.0, 0.0);
   hop_one<double2, MU, FWD>(sp, up, t);
#pragma unroll
   for (int c = 0; c < 12; c++) acc[c] = make_double2(acc[c].x - 0.5 * t[c].x, acc[c].y - 0.5 * t[c].y);
}

// the hops of direction MU whose neighbour lies inside (IN) or outside the
// block n (blk[y] == n); ifc selects one outside hop (-1: all inside ones)
template <int MU>
__device__ inline void aw_dir(const AwArgs &a, const int *blk, int n, int l, int x, int ifc, double2 *acc)
{
   const int yf = a.nbr[(size_t)(2 * MU) * a.V + x] & 0x7fffffff;
   const int yb = a.nbr[(size_t)(2 * MU + 1) * a.V + x] & 0x7fffffff;
   if (ifc < 0)
   {
      if (blk[yf] == n) aw_hop<MU, true>(a, l, x, yf, acc);
      if (blk[yb] == n) aw_hop<MU, false>(a, l, x, yb, acc);
   }
   else if (ifc == 2 * MU + 1)
   {
      if (blk[yf] != n) aw_hop<MU, true>(a, l, x, yf, acc);
   }
   else if (ifc == 2 * MU)
   {
      if (blk[yb] != n) aw_hop<MU, false>(a, l, x, yb, acc);
   }
}

// the hopping term of direction MU from outside the block n into x, if the
// neighbour in that direction lies in another block: its two colour
// products c0, c1 (times -1/2) and the face index (2 MU + 1 forward, 2 MU
// backward), else -1
template <int MU>
__device__ inline int aw_face(const AwArgs &a, const int *blk, int n, int l, int x, double2 *c0, double2 *c1)
{
   const int yf = a.nbr[(size_t)(2 * MU) * a.V + x] & 0x7fffffff;
   const int yb = a.nbr[(size_t)(2 * MU + 1) * a.V + x] & 0x7fffffff;
   const bool fwd = blk[yf] != n;
   if (!fwd && blk[yb] == n) return -1;
   const int y = fwd ? yf : yb;
   const double2 *up;
   if (fwd)
      up = (x >= a.h) ? link_ptr(a.ud, x - a.h, 2 * MU) : link_ptr(a.ud, y - a.h, 2 * MU + 1);
   else
      up = (x >= a.h) ? link_ptr(a.ud, x - a.h, 2 * MU + 1) : link_ptr(a.ud, y - a.h, 2 * MU);
   double2 sp[12], h0[3], h1[3];
   ld_basis64(a, l, y, sp);
   if (fwd)
   {
      spin_project<double2, MU, true>(sp, h0, h1);
      su3_apply<double2, false>(up, h0, c0);
      su3_apply<double2, false>(up, h1, c1);
   }
   else
   {
      spin_project<double2, MU, false>(sp, h0, h1);
      su3_apply<double2, true>(up, h0, c0);
      su3_apply<double2, true>(up, h1, c1);
   }
#pragma unroll
   for (int c = 0; c < 3; c++)
   {
      c0[c] = make_double2(-0.5 * c0[c].x, -0.5 * c0[c].y);
      c1[c] = make_double2(-0.5 * c1[c].x, -0.5 * c1[c].y);
   }
   return fwd ? 2 * MU + 1 : 2 * MU;
}

// (b, the spinor rebuilt from c0, c1 for direction MU) with b in FP32: the
// rebuild is the adjoint of the spin projection, so this equals
// (P b, c) with P b the projected components of b (in double precision)
template <int MU>
__device__ inline double2 aw_face_dot(const float2 *b, const double2 *c0, const double2 *c1, bool fwd)
{
   double2 bd[12], h0[3], h1[3];
#pragma unroll
   for (int c = 0; c < 12; c++) bd[c] = make_double2(b[c].x, b[c].y);
   if (fwd)
      spin_project<double2, MU, true>(bd, h0, h1);
   else
      spin_project<double2, MU, false>(bd, h0, h1);
   double2 z = make_double2(0.0, 0.0);
#pragma unroll
   for (int c = 0; c < 3; c++)
   {
      z.x += h0[c].x * c0[c].x + h0[c].y * c0[c].y + h1[c].x * c1[c].x + h1[c].y * c1[c].y;
      z.y += h0[c].x * c0[c].y - h0[c].y * c0[c].x + h1[c].x * c1[c].y - h1[c].y * c1[c].x;
   }
   return z;
}

// column l of Ablk and of the eight Ahop[ifc] of block n in one pass over
// the block's points: Y (site term and in-block hops of B_l, kept in LDS)
// and the up to four face terms of each point (in registers) are formed
// once; every basis vector B_k is then read once per point for both
__global__ __launch_bounds__(DWG) void k_aw_column(AwArgs a, const int *blk, int l)
{
   __shared__ double2 part[DWG / 64][9][NS_MAX];
   __shared__ double2 Ysh[12][DWG];  // Y of the lane's point (read back by the same lane)
   const int n = blockIdx.x;
   const int *pts = a.bpts + (size_t)n * a.bvol;
   const int Ns = a.Ns;
   const size_t nm = (size_t)Ns * Ns;
   const int w = threadIdx.x >> 6;
   const bool lead = (threadIdx.x & 63) == 0;
   for (int e = threadIdx.x; e < (DWG / 64) * 9 * NS_MAX; e += DWG)
      (&part[0][0][0])[e] = make_double2(0.0, 0.0);
   __syncthreads();
   for (int i0 = 0; i0 < a.bvol; i0 += DWG)
   {
      const int i = i0 + threadIdx.x;
      const int x = (i < a.bvol) ? pts[i] : 0;
      const bool act = (i < a.bvol) && !(a.masked && a.bnd[x]);
      if (!__any(act)) continue;
      // Y: site term and the hops from inside the block
      {
         double2 v[12];
#pragma unroll
         for (int c = 0; c < 12; c++) v[c] = make_double2(0.0, 0.0);
         if (act)
         {
            double2 s[12];
            ld_basis64(a, l, x, s);
            const double m = (x < a.h) ? a.mu : (a.eoflg ? 0.0 : a.mu);
            pauli_site(a.swd, x, m, s, v);
            aw_dir<0>(a, blk, n, l, x, -1, v);
            aw_dir<1>(a, blk, n, l, x, -1, v);
            aw_dir<2>(a, blk, n, l, x, -1, v);
            aw_dir<3>(a, blk, n, l, x, -1, v);
         }
#pragma unroll
         for (int c = 0; c < 12; c++) Ysh[c][threadIdx.x] = v[c];
      }
      // the face terms, then one pass over the basis vectors for both
      double2 f0[4][3], f1[4][3];
      int fs[4] = {-1, -1, -1, -1};
      if (act)
      {
         fs[0] = aw_face<0>(a, blk, n, l, x, f0[0], f1[0]);
         fs[1] = aw_face<1>(a, blk, n, l, x, f0[1], f1[1]);
         fs[2] = aw_face<2>(a, blk, n, l, x, f0[2], f1[2]);
         fs[3] = aw_face<3>(a, blk, n, l, x, f0[3], f1[3]);
      }
      bool anyf = false;
#pragma unroll
      for (int mu = 0; mu < 4; mu++) anyf = anyf || fs[mu] >= 0;
      const bool wf = __any(anyf);
      for (int k = 0; k < Ns; k++)
      {
         float2 b[12];
         if (act) load_spinor(a.basis + (size_t)k * a.fu, x, b);
         double2 d0 = make_double2(0.0, 0.0);
         if (act)
         {
            double2 v[12];
#pragma unroll
            for (int c = 0; c < 12; c++) v[c] = Ysh[c][threadIdx.x];
            d0 = sdot(b, v);
         }
         d0 = wave_red2(d0);
         if (lead)
         {
            part[w][0][k].x += d0.x;
            part[w][0][k].y += d0.y;
         }
         if (!wf) continue;
         double2 d[4];
         d[0] = (fs[0] >= 0) ? aw_face_dot<0>(b, f0[0], f1[0], fs[0] == 1) : make_double2(0.0, 0.0);
         d[1] = (fs[1] >= 0) ? aw_face_dot<1>(b, f0[1], f1[1], fs[1] == 3) : make_double2(0.0, 0.0);
         d[2] = (fs[2] >= 0) ? aw_face_dot<2>(b, f0[2], f1[2], fs[2] == 5) : make_double2(0.0, 0.0);
         d[3] = (fs[3] >= 0) ? aw_face_dot<3>(b, f0[3], f1[3], fs[3] == 7) : make_double2(0.0, 0.0);
#pragma unroll
         for (int ifc = 0; ifc < 8; ifc++)
         {
            const int mu = ifc >> 1;
            double2 e = (fs[mu] == ifc) ? d[mu] : make_double2(0.0, 0.0);
            e = wave_red2(e);
            if (lead)
            {
               part[w][1 + ifc][k].x += e.x;
               part[w][1 + ifc][k].y += e.y;
            }
         }
      }
   }
   __syncthreads();
   for (int e = threadIdx.x; e < 9 * Ns; e += DWG)
   {
      const int q = e / Ns, k = e % Ns;
      double2 s = make_double2(0.0, 0.0);
      for (int ww = 0; ww < DWG / 64; ww++)
      {
         s.x += part[ww][q][k].x;
         s.y += part[ww][q][k].y;
      }
      if (q == 0)
         a.Ablk[(size_t)n * nm + (size_t)k * Ns + l] = s;
      else
         a.Ahop[((size_t)n * 8 + (q - 1)) * nm + (size_t)k * Ns + l] = s;
   }
}

// per block: Ablk^-1 by Gauss-Jordan elimination with partial pivoting
// (double precision), the Frobenius condition number |A| |A^-1| against
// 1e6 (little/Aw_ops.c: MAX_FROBENIUS), then M[ifc] = Ablk^-1 Ahop[ifc];
// AinvT and MT are stored transposed ([column][row])
__global__ __launch_bounds__(64) void k_aw_invert(const double2 *Ablk, const double2 *Ahop, int Ns, double2 *AinvT,
                                                   double2 *MT, float2 *MT32, int *unsafe)
{
   extern __shared__ double2 lds[];
   double2 *A = lds;              // [Ns][2 Ns] augmented
   __shared__ int piv;
   __shared__ double fro[2];
   __shared__ double2 colc[NS_MAX];   // column c of the rows being eliminated
   const int n = blockIdx.x, t = threadIdx.x;
   const size_t nm = (size_t)Ns * Ns;
   const int W = 2 * Ns;
   for (int e = t; e < Ns * W; e += 64)
   {
      const int r = e / W, c = e % W;
      A[e] = (c < Ns) ? Ablk[(size_t)n * nm + (size_t)r * Ns + c]
                      : make_double2((c - Ns) == r ? 1.0 : 0.0, 0.0);
   }
   if (t == 0) fro[0] = fro[1] = 0.0;
   __syncthreads();
   if (t == 0)
   {
      double s = 0.0;
      for (int e = 0; e < Ns * W; e++)
         if (e % W < Ns) s += A[e].x * A[e].x + A[e].y * A[e].y;
      fro[0] = s;
   }
   for (int c = 0; c < Ns; c++)
   {
      __syncthreads();
      if (t == 0)
      {
         int p = c;
         double best = -1.0;
         for (int r = c; r < Ns; r++)
         {
            const double m = A[r * W + c].x * A[r * W + c].x + A[r * W + c].y * A[r * W + c].y;
            if (m > best) { best = m; p = r; }
         }
         piv = p;
      }
      __syncthreads();
      const int p = piv;
      if (p != c)
         for (int e = t; e < W; e += 64)
         {
            const double2 tmp = A[c * W + e];
            A[c * W + e] = A[p * W + e];
            A[p * W + e] = tmp;
         }
      __syncthreads();
      const double2 pv = A[c * W + c];
      const double pm = pv.x * pv.x + pv.y * pv.y;
      if (pm == 0.0)
      {
         if (t == 0) atomicOr(unsafe, 1);
         return;
      }
      const double2 ip = make_double2(pv.x / pm, -pv.y / pm);
      __syncthreads();
      for (int e = t; e < W; e += 64) A[c * W + e] = cmul(ip, A[c * W + e]);
      // the factors are read before any row is changed: a row of 2 Ns
      // entries may span two passes of the lanes over the matrix
      for (int r = t; r < Ns; r += 64) colc[r] = A[r * W + c];
      __syncthreads();
      for (int e = t; e < Ns * W; e += 64)
      {
         const int r = e / W, col = e % W;
         if (r == c) continue;
         A[e] = csub(A[e], cmul(colc[r], A[c * W + col]));
      }
   }
   __syncthreads();
   if (t == 0)
   {
      double s = 0.0;
      for (int r = 0; r < Ns; r++)
         for (int c = 0; c < Ns; c++)
         {
            const double2 v = A[r * W + Ns + c];
            s += v.x * v.x + v.y * v.y;
         }
      fro[1] = s;
      if (!(sqrt(fro[0] * fro[1]) < 1.0e6)) atomicOr(unsafe, 1);
   }
   // AinvT[c][r] = Ainv[r][c]
   for (int e = t; e < Ns * Ns; e += 64)
   {
      const int r = e / Ns, c = e % Ns;
      AinvT[(size_t)n * nm + (size_t)c * Ns + r] = A[r * W + Ns + c];
   }
   // M = Ainv Ahop[ifc], stored transposed: MT[ifc][l][k] = sum_j Ainv[k][j] Ahop[j][l]
   for (int f = 0; f < 8; f++)
   {
      const double2 *H = Ahop + ((size_t)n * 8 + f) * nm;
      for (int e = t; e < Ns * Ns; e += 64)
      {
         const int k = e / Ns, l = e % Ns;
         double2 s = make_double2(0.0, 0.0);
         for (int j = 0; j < Ns; j++) s = cadd(s, cmul(A[k * W + Ns + j], H[(size_t)j * Ns + l]));
         MT[((size_t)n * 8 + f) * nm + (size_t)l * Ns + k] = s;
         MT32[((size_t)n * 8 + f) * nm + (size_t)l * Ns + k] = make_float2((float)s.x, (float)s.y);
      }
   }
}

// ---------------------------------------------------------------------------
// the even-odd preconditioned little operator Ahat = 1 - Aee^-1 Aeo Aoo^-1 Aoe
// ---------------------------------------------------------------------------

// out[b][k] = sign * sum_ifc sum_l M[b0+b][ifc][k][l] in[inn[b0+b][ifc] - i0][l]
// (+ add[b][k] if add); one work-group of 256 lanes per block: lane group g
// (32 lanes for Ns <= 32, else 64) takes the faces g, g + 256/G, ..., lane k
// of a group the output component k; the face sums are added in LDS in a
// fixed order
constexpr int AWH = 256;
template <typename C>
__global__ __launch_bounds__(AWH) void k_aw_hop(const C *MT, const int *inn, int Ns, int b0, int i0, const C *in,
                                                 const C *add, C *out, int neg)
{
   typedef decltype(C::x) R;
   __shared__ C part[AWH];
   const int G = (Ns <= 32) ? 32 : 64, nf = AWH / G;
   const int b = blockIdx.x, g = threadIdx.x / G, k = threadIdx.x % G;
   const size_t nm = (size_t)Ns * Ns;
   const int n = b0 + b;
   R sx = 0, sy = 0;
   if (k < Ns)
      for (int f = g; f < 8; f += nf)
      {
         const C *M = MT + ((size_t)n * 8 + f) * nm;
         const C *v = in + (size_t)(inn[8 * n + f] - i0) * Ns;
         for (int l = 0; l < Ns; l++)
         {
            const C m = M[(size_t)l * Ns + k], x = v[l];
            sx += m.x * x.x - m.y * x.y;
            sy += m.x * x.y + m.y * x.x;
         }
      }
   part[threadIdx.x] = cmk<C>(sx, sy);
   __syncthreads();
   if (threadIdx.x >= Ns) return;
   sx = 0;
   sy = 0;
   for (int q = 0; q < nf; q++)
   {
      sx += part[q * G + k].x;
      sy += part[q * G + k].y;
   }
   if (neg)
   {
      sx = -sx;
      sy = -sy;
   }
   if (add)
   {
      sx += add[(size_t)b * Ns + k].x;
      sy += add[(size_t)b * Ns + k].y;
   }
   out[(size_t)b * Ns + k] = cmk<C>(sx, sy);
}

// the same for an even Ns <= 32: lane kk of a group forms the outputs
// 2 kk, 2 kk + 1 from adjacent entries of a row (one 16-byte load in FP32);
// group g takes the face g & 7 and the half g >> 3 of the sum over l, and
// the 16 partial sums of an output are added in LDS in a fixed order
template <typename C> struct Pair2
{
   decltype(C::x) ax, ay, bx, by;
};
template <typename C>
__global__ __launch_bounds__(AWH) void k_aw_hop2(const C *MT, const int *inn, int Ns, int b0, int i0, const C *in,
                                                 const C *add, C *out, int neg)
{
   typedef decltype(C::x) R;
   __shared__ Pair2<C> part[16][16];
   const int b = blockIdx.x, g = threadIdx.x >> 4, kk = threadIdx.x & 15;
   const int f = g & 7, hf = g >> 3;
   const size_t nm = (size_t)Ns * Ns;
   const int n = b0 + b;
   R s0x = 0, s0y = 0, s1x = 0, s1y = 0;
   if (2 * kk < Ns)
   {
      const C *M = MT + ((size_t)n * 8 + f) * nm;
      const C *v = in + (size_t)(inn[8 * n + f] - i0) * Ns;
      const int lh = Ns >> 1, l0 = hf ? lh : 0, l1 = hf ? Ns : lh;
      for (int l = l0; l < l1; l++)
      {
         C m0, m1;
         if constexpr (sizeof(C) == sizeof(float2))
         {
            const float4 m = *(const float4 *)(M + (size_t)l * Ns + 2 * kk);
            m0 = cmk<C>(m.x, m.y);
            m1 = cmk<C>(m.z, m.w);
         }
         else
         {
            m0 = M[(size_t)l * Ns + 2 * kk];
            m1 = M[(size_t)l * Ns + 2 * kk + 1];
         }
         const C x = v[l];
         s0x += m0.x * x.x - m0.y * x.y;
         s0y += m0.x * x.y + m0.y * x.x;
         s1x += m1.x * x.x - m1.y * x.y;
         s1y += m1.x * x.y + m1.y * x.x;
      }
   }
   part[g][kk] = {s0x, s0y, s1x, s1y};
   __syncthreads();
   const int k = threadIdx.x;
   if (k >= Ns) return;
   R sx = 0, sy = 0;
   for (int q = 0; q < 16; q++)
   {
      const Pair2<C> &p = part[q][k >> 1];
      sx += (k & 1) ? p.bx : p.ax;
      sy += (k & 1) ? p.by : p.ay;
   }
   if (neg)
   {
      sx = -sx;
      sy = -sy;
   }
   if (add)
   {
      sx += add[(size_t)b * Ns + k].x;
      sy += add[(size_t)b * Ns + k].y;
   }
   out[(size_t)b * Ns + k] = cmk<C>(sx, sy);
}

// out[b] = AinvT^T of block b0+b times in[b] (+ add)
__global__ __launch_bounds__(64) void k_aw_diag(const double2 *AinvT, int Ns, int b0, const double2 *in, double2 *out)
{
   const int b = blockIdx.x, k = threadIdx.x;
   if (k >= Ns) return;
   const size_t nm = (size_t)Ns * Ns;
   const double2 *A = AinvT + (size_t)(b0 + b) * nm;
   double2 s = make_double2(0.0, 0.0);
   for (int l = 0; l < Ns; l++) s = cadd(s, cmul(A[(size_t)l * Ns + k], in[(size_t)b * Ns + l]));
   out[(size_t)b * Ns + k] = s;
}

// Ahat v on the even blocks; t: work vector on the odd blocks
// one hop pass over the blocks [b0, b0 + nbh) (k_aw_hop2 for an even Ns <= 32)
template <typename C> void aw_hop(const C *MT, int b0, int i0, const C *in, const C *add, C *out, int neg)
{
   Lib &L = lib();
   if (D.Ns % 2 == 0 && D.Ns <= 32)
      k_aw_hop2<C><<<D.nbh, AWH, 0, L.stream>>>(MT, D.d_inn, D.Ns, b0, i0, in, add, out, neg);
   else
      k_aw_hop<C><<<D.nbh, AWH, 0, L.stream>>>(MT, D.d_inn, D.Ns, b0, i0, in, add, out, neg);
}

template <typename C> void ahat(const C *MT, const C *v, C *t, C *w)
{
   aw_hop<C>(MT, D.nbh, 0, v, nullptr, t, 0);
   aw_hop<C>(MT, 0, D.nbh, t, v, w, 1);
}

// ---------------------------------------------------------------------------
// little-vector linear algebra (contiguous complex vectors of n entries)
// ---------------------------------------------------------------------------

template <typename C>
__global__ __launch_bounds__(256) void k_vdot(const C *a, const C *b, size_t n, double *part)
{
   __shared__ double sh[3][256];
   double x = 0.0, y = 0.0, z = 0.0;
   for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
   {
      const double ax = a[i].x, ay = a[i].y, bx = b[i].x, by = b[i].y;
      x += ax * bx + ay * by;
      y += ax * by - ay * bx;
      z += bx * bx + by * by;
   }
   sh[0][threadIdx.x] = x;
   sh[1][threadIdx.x] = y;
   sh[2][threadIdx.x] = z;
   __syncthreads();
   for (int m = 128; m > 0; m >>= 1)
   {
      if (threadIdx.x < m)
         for (int q = 0; q < 3; q++) sh[q][threadIdx.x] += sh[q][threadIdx.x + m];
      __syncthreads();
   }
   if (threadIdx.x == 0)
      for (int q = 0; q < 3; q++) part[3 * blockIdx.x + q] = sh[q][0];
}

// (a, b) and |b|^2 in double, summed in a fixed order
template <typename C> void vdot(const C *a, const C *b, size_t n, double *re, double *im, double *nb2)
{
   Lib &L = lib();
   int nbk = (int)((n + 255) / 256);
   if (nbk > 1024) nbk = 1024;
   k_vdot<C><<<nbk, 256, 0, L.stream>>>(a, b, n, D.d_red);
   HIP_OK(hipMemcpyAsync(D.h_red, D.d_red, sizeof(double) * 3 * nbk, hipMemcpyDeviceToHost, L.stream));
   HIP_OK(hipStreamSynchronize(L.stream));
   double s[3] = {0.0, 0.0, 0.0};
   for (int b = 0; b < nbk; b++)
      for (int q = 0; q < 3; q++) s[q] += D.h_red[3 * b + q];
   if (re) *re = s[0];
   if (im) *im = s[1];
   if (nb2) *nb2 = s[2];
}

// y = y + z x (complex z)
template <typename C>
__global__ __launch_bounds__(256) void k_vaxpy(C *y, const C *x, size_t n, double zr, double zi)
{
   const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
   if (i >= n) return;
   typedef decltype(C::x) R;
   const R a = (R)zr, b = (R)zi;
   y[i] = cmk<C>(y[i].x + (a * x[i].x - b * x[i].y), y[i].y + (a * x[i].y + b * x[i].x));
}

template <typename C> void vaxpy(C *y, const C *x, size_t n, double zr, double zi)
{
   k_vaxpy<C><<<(unsigned)((n + 255) / 256), 256, 0, lib().stream>>>(y, x, n, zr, zi);
}

template <typename C>
__global__ __launch_bounds__(256) void k_vscale(C *y, size_t n, double s)
{
   const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
   if (i >= n) return;
   typedef decltype(C::x) R;
   y[i] = cmk<C>((R)s * y[i].x, (R)s * y[i].y);
}

template <typename A, typename B>
__global__ __launch_bounds__(256) void k_vconv(const A *a, B *b, size_t n)
{
   const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
   if (i >= n) return;
   typedef decltype(B::x) R;
   b[i] = cmk<B>((R)a[i].x, (R)a[i].y);
}

template <typename A, typename B> void vconv(const A *a, B *b, size_t n)
{
   k_vconv<A, B><<<(unsigned)((n + 255) / 256), 256, 0, lib().stream>>>(a, b, n);
}

double2 *wvd(int k)
{
   while ((int)D.wvd.size() <= k) D.wvd.push_back((double2 *)dev_alloc(sizeof(double2) * (size_t)D.nb * D.Ns));
   return D.wvd[k];
}

float2 *wv(int k)
{
   while ((int)D.wv.size() <= k) D.wv.push_back((float2 *)dev_alloc(sizeof(float2) * (size_t)D.nb * D.Ns));
   return D.wv[k];
}

// ---------------------------------------------------------------------------
// Gram-Schmidt of the little solvers with the coefficients kept on the device
// ---------------------------------------------------------------------------
//
// The Krylov vectors of a little solver are stored contiguously, phi_0..
// phi_{nmx-1} then chi_0..chi_{nmx-1}, ld complex numbers apart.  One step n
// of the classical Gram-Schmidt process of gcr4v.c / fgcr4vd.c runs as
//   k_lmdot + k_lsum:      a[n][l] = (chi_l, chi_n), l < n
//   k_lmaxpy + k_lsum_bc:  chi_n -= sum_l a[n][l] chi_l; b_n = |chi_n|,
//                          c_n = (chi_n, rho) / b_n
//   k_lscale + k_lsum_bc:  chi_n /= b_n; rho -= c_n chi_n; |rho|^2
// and the host reads the four numbers of step n once.  The sums are taken in
// double precision in a fixed order (per lane, per work-group, then over the
// work-groups), so the results do not depend on the launch timing.

constexpr int LRB = 256;     // lanes of the little-vector kernels
constexpr int LNBK = 512;    // largest number of work-groups of a reduction

int lnbk(size_t nv)
{
   size_t n = (nv + 4 * LRB - 1) / (4 * LRB);
   return (int)std::max<size_t>(1, std::min<size_t>(n, LNBK));
}

__device__ inline double2 wave_sum2(double2 v)
{
#pragma unroll
   for (int m = 32; m >= 1; m >>= 1)
   {
      v.x += __shfl_xor(v.x, m);
      v.y += __shfl_xor(v.y, m);
   }
   return v;
}

// part[l * gridDim.x + blockIdx.x] = this work-group's share of (X_l, y)
template <typename C>
__global__ __launch_bounds__(LRB) void k_lmdot(const C *X, size_t ld, const C *yv, size_t nv, double2 *part)
{
   __shared__ double2 sh[LRB / 64];
   const int l = blockIdx.y;
   const C *cl = X + (size_t)l * ld, *cn = yv;
   double x = 0.0, y = 0.0;
   for (size_t i = (size_t)blockIdx.x * LRB + threadIdx.x; i < nv; i += (size_t)gridDim.x * LRB)
   {
      const double ax = cl[i].x, ay = cl[i].y, bx = cn[i].x, by = cn[i].y;
      x += ax * bx + ay * by;
      y += ax * by - ay * bx;
   }
   const double2 s = wg_sum(make_double2(x, y), sh);
   if (threadIdx.x == 0) part[(size_t)l * gridDim.x + blockIdx.x] = s;
}

// out[q] = sum_b part[q * nbk + b] for q < nq; one wave per q, fixed order
__global__ __launch_bounds__(LRB) void k_lsum(const double2 *part, int nbk, int nq, double2 *out)
{
   const int w = threadIdx.x >> 6, t = threadIdx.x & 63;
   for (int q = w; q < nq; q += LRB / 64)
   {
      double2 v = make_double2(0.0, 0.0);
      for (int b = t; b < nbk; b += 64)
      {
         v.x += part[(size_t)q * nbk + b].x;
         v.y += part[(size_t)q * nbk + b].y;
      }
      v = wave_sum2(v);
      if (t == 0) out[q] = v;
   }
}

// mode 0: bc[0] = |chi_n|, bc[1] + i bc[2] = (chi_n, rho) / |chi_n| from the
// columns (chi_n, rho) and (|chi_n|^2, 0); mode 1: bc[3] = |rho|^2
__global__ __launch_bounds__(64) void k_lsum_bc(const double2 *part, int nbk, int mode, double *bc)
{
   const int t = threadIdx.x;
   double2 v0 = make_double2(0.0, 0.0), v1 = make_double2(0.0, 0.0);
   for (int b = t; b < nbk; b += 64)
   {
      v0.x += part[b].x;
      v0.y += part[b].y;
      if (mode == 0) v1.x += part[nbk + b].x;
   }
   v0 = wave_sum2(v0);
   v1 = wave_sum2(v1);
   if (t != 0) return;
   if (mode == 0)
   {
      const double bn = sqrt(v1.x);
      bc[0] = bn;
      bc[1] = (bn > 0.0) ? v0.x / bn : 0.0;
      bc[2] = (bn > 0.0) ? v0.y / bn : 0.0;
   }
   else
      bc[3] = v0.x;
}

// chi_n -= sum_{l<n} a[l] chi_l (in the precision of C), then this
// work-group's shares of (chi_n, rho) and |chi_n|^2
template <typename C>
__global__ __launch_bounds__(LRB) void k_lmaxpy(C *chi, size_t ld, int n, size_t nv, const double2 *a, const C *rho,
                                                double2 *part)
{
   typedef decltype(C::x) R;
   __shared__ double2 sh[LRB / 64];
   C *cn = chi + (size_t)n * ld;
   double x = 0.0, y = 0.0, z = 0.0;
   for (size_t i = (size_t)blockIdx.x * LRB + threadIdx.x; i < nv; i += (size_t)gridDim.x * LRB)
   {
      R vx = cn[i].x, vy = cn[i].y;
      for (int l = 0; l < n; l++)
      {
         const R ar = (R)a[l].x, ai = (R)a[l].y;
         const C w = chi[(size_t)l * ld + i];
         vx -= ar * w.x - ai * w.y;
         vy -= ar * w.y + ai * w.x;
      }
      cn[i] = cmk<C>(vx, vy);
      const double px = vx, py = vy, rx = rho[i].x, ry = rho[i].y;
      x += px * rx + py * ry;
      y += px * ry - py * rx;
      z += px * px + py * py;
   }
   const double2 s0 = wg_sum(make_double2(x, y), sh);
   const double2 s1 = wg_sum(make_double2(z, 0.0), sh);
   if (threadIdx.x == 0)
   {
      part[blockIdx.x] = s0;
      part[gridDim.x + blockIdx.x] = s1;
   }
}

// chi_n /= b_n, rho -= c_n chi_n (nothing if b_n = 0), shares of |rho|^2
template <typename C>
__global__ __launch_bounds__(LRB) void k_lscale(C *cn, C *rho, size_t nv, const double *bc, double2 *part)
{
   typedef decltype(C::x) R;
   __shared__ double2 sh[LRB / 64];
   const double bn = bc[0];
   const R sc = (bn > 0.0) ? (R)(1.0 / bn) : (R)0, cr = (R)bc[1], ci = (R)bc[2];
   double z = 0.0;
   for (size_t i = (size_t)blockIdx.x * LRB + threadIdx.x; i < nv; i += (size_t)gridDim.x * LRB)
   {
      R rx = rho[i].x, ry = rho[i].y;
      if (bn > 0.0)
      {
         const R vx = sc * cn[i].x, vy = sc * cn[i].y;
         cn[i] = cmk<C>(vx, vy);
         rx -= cr * vx - ci * vy;
         ry -= cr * vy + ci * vx;
         rho[i] = cmk<C>(rx, ry);
      }
      const double px = rx, py = ry;
      z += px * px + py * py;
   }
   const double2 s = wg_sum(make_double2(z, 0.0), sh);
   if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// x = (add ? x : 0) + sum_{l<m} c[l] phi_l
template <typename C>
__global__ __launch_bounds__(LRB) void k_lcomb(C *x, const C *phi, size_t ld, int m, size_t nv, const double2 *c,
                                               int add)
{
   typedef decltype(C::x) R;
   const size_t i = (size_t)blockIdx.x * LRB + threadIdx.x;
   if (i >= nv) return;
   R vx = 0, vy = 0;
   if (add)
   {
      vx = x[i].x;
      vy = x[i].y;
   }
   for (int l = 0; l < m; l++)
   {
      const R cr = (R)c[l].x, ci = (R)c[l].y;
      const C w = phi[(size_t)l * ld + i];
      vx += cr * w.x - ci * w.y;
      vy += cr * w.y + ci * w.x;
   }
   x[i] = cmk<C>(vx, vy);
}

// one little solver's Krylov space and Gram-Schmidt state
template <typename C> struct LKry
{
   C *v = nullptr;          // [2 nmx][ld]: phi_0.., then chi_0..
   int nmx = 0;
   size_t ld = 0, nv = 0;
   double2 *a = nullptr;    // [nmx][nmx]: a[n * nmx + l] = (chi_l, chi_n)
   double *bc = nullptr;    // [nmx][4]: |chi_n|, c_n, |rho|^2 after step n
   double2 *c = nullptr;    // [nmx] coefficients of the solution
   double2 *part = nullptr; // [nmx][LNBK] work-group shares
   double *h = nullptr;     // pinned host copy of a, bc, then c
   C *phi(int k) const { return v + (size_t)k * ld; }
   C *chi(int k) const { return v + (size_t)(nmx + k) * ld; }
};

LKry<float2> K32;
LKry<double2> K64;

template <typename C> void lkry_free(LKry<C> &K)
{
   dev_free(K.v); dev_free(K.a); dev_free(K.bc); dev_free(K.c); dev_free(K.part);
   if (K.h) (void)hipHostFree(K.h);
   K = LKry<C>();
}

template <typename C> bool lkry_alloc(LKry<C> &K, int nmx, size_t nv)
{
   const size_t ld = (nv + 63) & ~(size_t)63;
   if (K.v && K.nmx == nmx && K.ld == ld) return true;
   lkry_free(K);
   const int nm = std::max(nmx, 2);
   K.v = (C *)dev_alloc(sizeof(C) * 2 * (size_t)nmx * ld);
   K.a = (double2 *)dev_alloc(sizeof(double2) * (size_t)nmx * nmx);
   K.bc = (double *)dev_alloc(sizeof(double) * 4 * (size_t)nmx);
   K.c = (double2 *)dev_alloc(sizeof(double2) * (size_t)nmx);
   K.part = (double2 *)dev_alloc(sizeof(double2) * (size_t)nm * LNBK);
   HIP_OK(hipHostMalloc((void **)&K.h, sizeof(double) * (2 * (size_t)nmx * nmx + 6 * (size_t)nmx), 0));
   if (!K.v || !K.a || !K.bc || !K.c || !K.part || !K.h)
   {
      lkry_free(K);
      return false;
   }
   K.nmx = nmx;
   K.ld = ld;
   K.nv = nv;
   return true;
}

void dfl_free_grid()
{
   dev_free(D.d_bpts); dev_free(D.d_inn); dev_free(D.d_blk); dev_free(D.d_basis); dev_free(D.d_vmds); dev_free(D.d_s64);
   dev_free(D.d_Ablk); dev_free(D.d_Ahop); dev_free(D.d_AinvT); dev_free(D.d_MT); dev_free(D.d_MT32);
   dev_free(D.d_unsafe); dev_free(D.d_red);
   for (auto p : D.wvd) dev_free(p);
   for (auto p : D.wv) dev_free(p);
   if (D.h_red) (void)hipHostFree(D.h_red);
   D.d_bpts = D.d_inn = D.d_blk = nullptr;
   D.d_basis = nullptr; D.d_vmds = nullptr; D.d_s64 = nullptr;
   D.d_Ablk = D.d_Ahop = D.d_AinvT = D.d_MT = nullptr;
   D.d_MT32 = nullptr; D.d_unsafe = nullptr; D.d_red = nullptr; D.h_red = nullptr;
   D.wvd.clear(); D.wv.clear();
   D.grid = D.basis_ok = D.aw_ok = D.awhat_ok = false;
   lkry_free(K32);
   lkry_free(K64);
   dev_free(D.d_lw); dev_free(D.d_law); dev_free(D.d_lw32); dev_free(D.d_law32); dev_free(D.d_lM);
   dev_free(D.d_lc); dev_free(D.d_ld); dev_free(D.d_lB); dev_free(D.d_lpart); dev_free(D.d_lbc);
   D.d_lw = D.d_law = nullptr; D.d_lw32 = D.d_law32 = nullptr; D.d_lM = D.d_lc = D.d_ld = D.d_lB = nullptr;
   D.d_lpart = nullptr; D.d_lbc = nullptr;
   D.lm_ok = false;
   dev_free(D.d_gcoef); dev_free(D.d_ggram); dev_free(D.d_gscr);
   D.d_gcoef = D.d_ggram = D.d_gscr = nullptr;
   D.gcap = 0;
}

// step n of the Gram-Schmidt process on chi_n (see above); returns false if
// |chi_n| = 0, else the residual norm |rho| after the step in *rn
template <typename C> bool lgs_step(LKry<C> &K, int n, C *rho, double *rn)
{
   Lib &L = lib();
   const size_t nv = K.nv;
   const int nbk = lnbk(nv);
   double2 *acol = K.a + (size_t)n * K.nmx;
   double *bcn = K.bc + 4 * (size_t)n;
   if (n > 0)
   {
      k_lmdot<C><<<dim3(nbk, n), LRB, 0, L.stream>>>(K.chi(0), K.ld, K.chi(n), nv, K.part);
      k_lsum<<<1, LRB, 0, L.stream>>>(K.part, nbk, n, acol);
   }
   k_lmaxpy<C><<<nbk, LRB, 0, L.stream>>>(K.chi(0), K.ld, n, nv, acol, rho, K.part);
   k_lsum_bc<<<1, 64, 0, L.stream>>>(K.part, nbk, 0, bcn);
   k_lscale<C><<<nbk, LRB, 0, L.stream>>>(K.chi(n), rho, nv, bcn, K.part);
   k_lsum_bc<<<1, 64, 0, L.stream>>>(K.part, nbk, 1, bcn);
   HIP_OK(hipMemcpyAsync(K.h, bcn, sizeof(double) * 4, hipMemcpyDeviceToHost, L.stream));
   HIP_OK(hipStreamSynchronize(L.stream));
   *rn = sqrt(K.h[3]);
   return K.h[0] > 0.0;
}

// x = (add ? x : 0) + sum_{l<m} c'_l phi_l with c' from the back
// substitution of gcr4v.c / fgcr4vd.c (update_psi)
template <typename C> void lgs_solution(LKry<C> &K, int m, C *x, int add)
{
   Lib &L = lib();
   if (m <= 0)
   {
      if (!add) HIP_OK(hipMemsetAsync(x, 0, sizeof(C) * K.nv, L.stream));
      return;
   }
   const int nmx = K.nmx;
   double *ha = K.h, *hb = K.h + 2 * (size_t)nmx * nmx, *hc = hb + 4 * (size_t)nmx;
   HIP_OK(hipMemcpyAsync(ha, K.a, sizeof(double2) * (size_t)m * nmx, hipMemcpyDeviceToHost, L.stream));
   HIP_OK(hipMemcpyAsync(hb, K.bc, sizeof(double) * 4 * (size_t)m, hipMemcpyDeviceToHost, L.stream));
   HIP_OK(hipStreamSynchronize(L.stream));
   for (int l = m - 1; l >= 0; l--)
   {
      double zr = hb[4 * l + 1], zi = hb[4 * l + 2];
      for (int i = l + 1; i < m; i++)
      {
         // (chi_l, chi_i) is stored in column i
         const double ar = ha[2 * ((size_t)i * nmx + l)], ai = ha[2 * ((size_t)i * nmx + l) + 1];
         const double cr = hc[2 * i], ci = hc[2 * i + 1];
         zr -= ar * cr - ai * ci;
         zi -= ar * ci + ai * cr;
      }
      hc[2 * l] = zr / hb[4 * l];
      hc[2 * l + 1] = zi / hb[4 * l];
   }
   HIP_OK(hipMemcpyAsync(K.c, hc, sizeof(double2) * (size_t)m, hipMemcpyHostToDevice, L.stream));
   k_lcomb<C><<<(unsigned)((K.nv + LRB - 1) / LRB), LRB, 0, L.stream>>>(x, K.phi(0), K.ld, m, K.nv, K.c, add);
   // the pinned coefficients must stay until the copy has run
   HIP_OK(hipStreamSynchronize(L.stream));
}

// ---------------------------------------------------------------------------
// deflation of the little equation by the little modes (little/ltl_modes.c)
// ---------------------------------------------------------------------------
//
// The little modes w_l are the even-block parts of the coordinates of the Ns
// generating modes in the subspace basis (dfl_subspace), orthonormalised by
// classical Gram-Schmidt.  With P0 the projector to their span and
// M = (P0 Ahat P0)^-1 (the Ns x Ns matrix (w_l, Ahat w_j)^-1), the little
// solver works with P_L = 1 - Ahat P0 M P0, which removes the lowest modes of
// Ahat from the Krylov iteration.  All coefficients stay on the device.

// d[k] = sign * sum_l M[k][l] c[l], and d[Ns + k] = -d[k]
__global__ __launch_bounds__(64) void k_lmatvec(const double2 *M, int Ns, const double2 *c, double sign, double2 *d)
{
   const int k = threadIdx.x;
   if (k >= Ns) return;
   double2 s = make_double2(0.0, 0.0);
   for (int l = 0; l < Ns; l++) s = cadd(s, cmul(M[(size_t)k * Ns + l], c[l]));
   d[k] = make_double2(sign * s.x, sign * s.y);
   d[Ns + k] = make_double2(-sign * s.x, -sign * s.y);
}

// v /= bc[0] (bc[0] > 0)
template <typename C> __global__ __launch_bounds__(LRB) void k_lnorm(C *v, size_t nv, const double *bc)
{
   typedef decltype(C::x) R;
   const size_t i = (size_t)blockIdx.x * LRB + threadIdx.x;
   if (i >= nv || !(bc[0] > 0.0)) return;
   const R f = (R)(1.0 / bc[0]);
   v[i] = cmk<C>(f * v[i].x, f * v[i].y);
}

// D.d_ld = sign * M (X, v), and its negative at D.d_ld + Ns
template <typename C> void lm_coef(const C *X, const C *v, double sign)
{
   Lib &L = lib();
   const size_t nv = (size_t)D.nbh * D.Ns;
   const int nbk = lnbk(nv);
   k_lmdot<C><<<dim3(nbk, D.Ns), LRB, 0, L.stream>>>(X, D.ldv, v, nv, D.d_lpart);
   k_lsum<<<1, LRB, 0, L.stream>>>(D.d_lpart, nbk, D.Ns, D.d_lc);
   k_lmatvec<<<1, 64, 0, L.stream>>>(D.d_lM, D.Ns, D.d_lc, sign, D.d_ld);
}

template <typename C> void lm_comb(C *x, const C *Y, const double2 *c, int add)
{
   const size_t nv = (size_t)D.nbh * D.Ns;
   k_lcomb<C><<<(unsigned)((nv + LRB - 1) / LRB), LRB, 0, lib().stream>>>(x, Y, D.ldv, D.Ns, nv, c, add);
}

// vd = P_L vd (dfl_Lvd)
void dfl_Lvd_dev(double2 *vd)
{
   lm_coef<double2>(D.d_lw, vd, 1.0);
   lm_comb<double2>(vd, D.d_law, D.d_ld + D.Ns, 1);
}

// wd = P0 M P0 vd, vd = P_L vd (dfl_LRvd)
void dfl_LRvd_dev(double2 *vd, double2 *wd)
{
   lm_coef<double2>(D.d_lw, vd, 1.0);
   lm_comb<double2>(wd, D.d_lw, D.d_ld, 0);
   lm_comb<double2>(vd, D.d_law, D.d_ld + D.Ns, 1);
}

// vd -= P0 M P0 wd, wd = P_L wd (dfl_RLvd)
void dfl_RLvd_dev(double2 *vd, double2 *wd)
{
   lm_coef<double2>(D.d_lw, wd, -1.0);
   lm_comb<double2>(wd, D.d_law, D.d_ld, 1);
   lm_comb<double2>(vd, D.d_lw, D.d_ld, 1);
}

// v = P_L v in single precision (dfl_Lv)
void dfl_Lv_dev(float2 *v)
{
   lm_coef<float2>(D.d_lw32, v, 1.0);
   lm_comb<float2>(v, D.d_law32, D.d_ld + D.Ns, 1);
}

inline double2 hmul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
inline double2 hsub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }

// the little modes, Ahat on them and M (set_ltl_modes); returns 1 if the
// inversion of (w, Ahat w) is not safe (Frobenius condition number >= 1e6)
// or the modes are degenerate, else 0
int set_ltl_modes()
{
   Lib &L = lib();
   const int Ns = D.Ns;
   const size_t nv = (size_t)D.nbh * Ns;
   const size_t ldv = (nv + 63) & ~(size_t)63;
   const int nbk = lnbk(nv);
   if (!D.d_lw || D.ldv != ldv)
   {
      dev_free(D.d_lw); dev_free(D.d_law); dev_free(D.d_lw32); dev_free(D.d_law32); dev_free(D.d_lM);
      dev_free(D.d_lc); dev_free(D.d_ld); dev_free(D.d_lB); dev_free(D.d_lpart); dev_free(D.d_lbc);
      D.ldv = ldv;
      D.d_lw = (double2 *)dev_alloc(sizeof(double2) * ldv * Ns);
      D.d_law = (double2 *)dev_alloc(sizeof(double2) * ldv * Ns);
      D.d_lw32 = (float2 *)dev_alloc(sizeof(float2) * ldv * Ns);
      D.d_law32 = (float2 *)dev_alloc(sizeof(float2) * ldv * Ns);
      D.d_lM = (double2 *)dev_alloc(sizeof(double2) * Ns * Ns);
      D.d_lc = (double2 *)dev_alloc(sizeof(double2) * Ns);
      D.d_ld = (double2 *)dev_alloc(sizeof(double2) * 2 * Ns);
      D.d_lB = (double2 *)dev_alloc(sizeof(double2) * Ns * Ns);
      D.d_lpart = (double2 *)dev_alloc(sizeof(double2) * (size_t)std::max(Ns, 2) * LNBK);
      D.d_lbc = (double *)dev_alloc(sizeof(double) * 4 * Ns);
      if (!D.d_lw || !D.d_law || !D.d_lw32 || !D.d_law32 || !D.d_lM || !D.d_lc || !D.d_ld || !D.d_lB ||
          !D.d_lpart || !D.d_lbc)
         return 1;
   }
   D.lm_ok = false;
   // the modes' coordinates on the even blocks, Gram-Schmidt (vs2vds)
   double2 *acol = D.d_lB;
   for (int l = 0; l < Ns; l++)
   {
      double2 *wl = D.d_lw + (size_t)l * ldv;
      vconv<float2, double2>(D.d_vmds + (size_t)l * D.nb * Ns, wl, nv);
      if (l > 0)
      {
         k_lmdot<double2><<<dim3(nbk, l), LRB, 0, L.stream>>>(D.d_lw, ldv, wl, nv, D.d_lpart);
         k_lsum<<<1, LRB, 0, L.stream>>>(D.d_lpart, nbk, l, acol);
      }
      k_lmaxpy<double2><<<nbk, LRB, 0, L.stream>>>(D.d_lw, ldv, l, nv, acol, wl, D.d_lpart);
      k_lsum_bc<<<1, 64, 0, L.stream>>>(D.d_lpart, nbk, 0, D.d_lbc + 4 * l);
      k_lnorm<double2><<<(unsigned)((nv + LRB - 1) / LRB), LRB, 0, L.stream>>>(wl, nv, D.d_lbc + 4 * l);
   }
   // Ahat w_l (vds2Avds) and the FP32 copies
   double2 *t = wvd(8);
   for (int l = 0; l < Ns; l++)
   {
      ahat<double2>(D.d_MT, D.d_lw + (size_t)l * ldv, t, D.d_law + (size_t)l * ldv);
      vconv<double2, float2>(D.d_lw + (size_t)l * ldv, D.d_lw32 + (size_t)l * ldv, nv);
      vconv<double2, float2>(D.d_law + (size_t)l * ldv, D.d_law32 + (size_t)l * ldv, nv);
   }
   // B[l][j] = (w_l, Ahat w_j), stored by columns (set_Ads)
   for (int j = 0; j < Ns; j++)
   {
      k_lmdot<double2><<<dim3(nbk, Ns), LRB, 0, L.stream>>>(D.d_lw, ldv, D.d_law + (size_t)j * ldv, nv, D.d_lpart);
      k_lsum<<<1, LRB, 0, L.stream>>>(D.d_lpart, nbk, Ns, D.d_lB + (size_t)j * Ns);
   }
   std::vector<double2> Bc((size_t)Ns * Ns);
   std::vector<double> bc(4 * (size_t)Ns);
   HIP_OK(hipMemcpyAsync(Bc.data(), D.d_lB, sizeof(double2) * Ns * Ns, hipMemcpyDeviceToHost, L.stream));
   HIP_OK(hipMemcpyAsync(bc.data(), D.d_lbc, sizeof(double) * 4 * Ns, hipMemcpyDeviceToHost, L.stream));
   HIP_OK(hipStreamSynchronize(L.stream));
   for (int l = 0; l < Ns; l++)
      if (!(bc[4 * l] > 0.0)) return 1;
   // M = B^-1: Gauss-Jordan with partial pivoting in double precision
   const int W = 2 * Ns;
   std::vector<double2> A((size_t)Ns * W);
   double f0 = 0.0, f1 = 0.0;
   for (int r = 0; r < Ns; r++)
      for (int c = 0; c < W; c++)
      {
         A[(size_t)r * W + c] = (c < Ns) ? Bc[(size_t)c * Ns + r] : make_double2(c - Ns == r ? 1.0 : 0.0, 0.0);
         if (c < Ns) f0 += A[(size_t)r * W + c].x * A[(size_t)r * W + c].x + A[(size_t)r * W + c].y * A[(size_t)r * W + c].y;
      }
   for (int c = 0; c < Ns; c++)
   {
      int p = c;
      double best = -1.0;
      for (int r = c; r < Ns; r++)
      {
         const double m = A[(size_t)r * W + c].x * A[(size_t)r * W + c].x + A[(size_t)r * W + c].y * A[(size_t)r * W + c].y;
         if (m > best) { best = m; p = r; }
      }
      if (!(best > 0.0)) return 1;
      if (p != c)
         for (int e = 0; e < W; e++) std::swap(A[(size_t)c * W + e], A[(size_t)p * W + e]);
      const double2 pv = A[(size_t)c * W + c];
      const double pm = pv.x * pv.x + pv.y * pv.y;
      const double2 ip = make_double2(pv.x / pm, -pv.y / pm);
      for (int e = 0; e < W; e++) A[(size_t)c * W + e] = hmul(ip, A[(size_t)c * W + e]);
      for (int r = 0; r < Ns; r++)
      {
         if (r == c) continue;
         const double2 fr = A[(size_t)r * W + c];
         for (int e = 0; e < W; e++) A[(size_t)r * W + e] = hsub(A[(size_t)r * W + e], hmul(fr, A[(size_t)c * W + e]));
      }
   }
   std::vector<double2> Mi((size_t)Ns * Ns);
   for (int r = 0; r < Ns; r++)
      for (int c = 0; c < Ns; c++)
      {
         Mi[(size_t)r * Ns + c] = A[(size_t)r * W + Ns + c];
         f1 += Mi[(size_t)r * Ns + c].x * Mi[(size_t)r * Ns + c].x + Mi[(size_t)r * Ns + c].y * Mi[(size_t)r * Ns + c].y;
      }
   if (!(sqrt(f0 * f1) < 1.0e6)) return 1;
   HIP_OK(hipMemcpyAsync(D.d_lM, Mi.data(), sizeof(double2) * Ns * Ns, hipMemcpyHostToDevice, L.stream));
   HIP_OK(hipStreamSynchronize(L.stream));
   D.lm_ok = true;
   return 0;
}

// single-pass GCR in single precision for Ahat x = b (linsolv/gcr4v.c):
// returns the number of Krylov vectors (or -1 if res was not reached in nmx)
int gcr4v(int nmx, double res, const float2 *b, float2 *x)
{
   const size_t nv = (size_t)D.nbh * D.Ns;
   if (!lkry_alloc(K32, nmx, nv)) return -1;
   float2 *rho = wv(0), *t = wv(1);
   Lib &L = lib();
   HIP_OK(hipMemcpyAsync(rho, b, sizeof(float2) * nv, hipMemcpyDeviceToDevice, L.stream));
   double r0;
   vdot<float2>(rho, rho, nv, nullptr, nullptr, &r0);
   r0 = sqrt(r0);
   double rn = r0;
   int n = 0, status = -1;
   if (r0 > 0.0)
      for (; n < nmx; n++)
      {
         ahat<float2>(D.d_MT32, rho, t, K32.chi(n));
         dfl_Lv_dev(K32.chi(n));
         HIP_OK(hipMemcpyAsync(K32.phi(n), rho, sizeof(float2) * nv, hipMemcpyDeviceToDevice, L.stream));
         if (!lgs_step(K32, n, rho, &rn)) break;
         if (rn <= res * r0)
         {
            n++;
            status = n;
            break;
         }
      }
   else
      status = 0;
   lgs_solution(K32, n, x, 0);
   return status;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI: parameters (flags/dfl_parms.c)
// ---------------------------------------------------------------------------

extern "C" dfl_parms_t set_dfl_parms(int *bs, int Ns)
{
   const char *where = "set_dfl_parms [dfl_parms.c]";
   Lib &L = lib();
   dfl_parms_t p;
   memset(&p, 0, sizeof(p));
   if (!require_geometry(where)) return p;
   if (D.Ns > 0 && (memcmp(bs, D.bs, sizeof(D.bs)) != 0 || Ns != D.Ns))
   {
      fail(where, "bs[4] and Ns may be set only once");
      return p;
   }
   int nblk = 1;
   for (int mu = 0; mu < 4; mu++)
   {
      if (bs[mu] < 4 || bs[mu] > L.g.L[mu] || (bs[mu] % 2) || (L.g.L[mu] % bs[mu]) || ((L.g.L[mu] / bs[mu]) % 2))
      {
         fail(where, "Block sizes are out of range, odd, do not divide the lattice or give an odd number of blocks");
         return p;
      }
      nblk *= L.g.L[mu] / bs[mu];
   }
   if (nblk % (2 * L.g.nthread))
   {
      fail(where, "The number of blocks in the local lattice must be an even multiple of NTHREAD");
      return p;
   }
   if (Ns < 2 || (Ns & 1) || Ns > NS_MAX)
   {
      fail(where, "Improper value of Ns");
      return p;
   }
   memcpy(D.bs, bs, sizeof(D.bs));
   D.Ns = Ns;
   memcpy(p.bs, bs, sizeof(p.bs));
   p.Ns = Ns;
   return p;
}

extern "C" dfl_parms_t dfl_parms(void)
{
   dfl_parms_t p;
   memcpy(p.bs, D.bs, sizeof(p.bs));
   p.Ns = D.Ns;
   return p;
}

extern "C" dfl_pro_parms_t set_dfl_pro_parms(int nkv, int nmx, double res, int nmx_gcr, double res_gcr)
{
   dfl_pro_parms_t p;
   memset(&p, 0, sizeof(p));
   if (nkv < 1 || nmx < 1 || res <= DBL_EPSILON || nmx_gcr < 2 || res_gcr <= FLT_EPSILON)
   {
      fail("set_dfl_pro_parms [dfl_parms.c]", "Parameters are out of range");
      return p;
   }
   D.nkv = nkv;
   D.nmx = nmx;
   D.res = res;
   D.nmx_gcr = nmx_gcr;
   D.res_gcr = res_gcr;
   p.nkv = nkv;
   p.nmx = nmx;
   p.nmx_gcr = nmx_gcr;
   p.res = res;
   p.res_gcr = res_gcr;
   return p;
}

extern "C" dfl_pro_parms_t dfl_pro_parms(void)
{
   dfl_pro_parms_t p;
   p.nkv = D.nkv;
   p.nmx = D.nmx;
   p.nmx_gcr = D.nmx_gcr;
   p.res = D.res;
   p.res_gcr = D.res_gcr;
   return p;
}

extern "C" dfl_gen_parms_t set_dfl_gen_parms(double kappa, double mu, int ninv, int nmr, int ncy)
{
   dfl_gen_parms_t p;
   memset(&p, 0, sizeof(p));
   if (ninv < 4 || nmr < 1 || ncy < 1 || kappa < 0.0)
   {
      fail("set_dfl_gen_parms [dfl_parms.c]", "Parameters are out of range");
      return p;
   }
   D.ninv = ninv;
   D.nmr = nmr;
   D.ncy = ncy;
   D.kappa = kappa;
   D.mu = mu;
   D.m0 = (kappa != 0.0) ? 1.0 / (2.0 * kappa) - 4.0 : DBL_MAX;
   p.ninv = ninv;
   p.nmr = nmr;
   p.ncy = ncy;
   p.kappa = kappa;
   p.m0 = D.m0;
   p.mu = mu;
   return p;
}

extern "C" dfl_gen_parms_t dfl_gen_parms(void)
{
   dfl_gen_parms_t p;
   p.ninv = D.ninv;
   p.nmr = D.nmr;
   p.ncy = D.ncy;
   p.kappa = D.kappa;
   p.m0 = D.m0;
   p.mu = D.mu;
   return p;
}

// ---------------------------------------------------------------------------
// C ABI: the deflation subspace (dfl/dfl_subspace.c)
// ---------------------------------------------------------------------------

namespace {

// device little vector <-> host array of nb * Ns complex_dble
bool s2vd_dev(const float2 *s, double2 *vd)
{
   Lib &L = lib();
   if (D.bvol <= DWG)
      k_blk_proj_reg<float2><<<D.nb, DWG, 0, L.stream>>>(D.d_basis, D.fu, D.Ns, D.d_bpts, D.bvol, s, vd);
   else
      k_blk_proj<float2><<<D.nb, DWG, 0, L.stream>>>(D.d_basis, D.fu, D.Ns, D.d_bpts, D.bvol, s, vd);
   return true;
}

void vd2s_dev(const double2 *vd, float2 *s)
{
   Lib &L = lib();
   k_blk_prolong<<<D.nb, DWG, 0, L.stream>>>(D.d_basis, D.fu, D.Ns, D.d_bpts, D.bvol, vd, s);
}

bool ready(const char *where)
{
   if (!dfl_grid(where)) return false;
   if (!D.basis_ok)
   {
      fail(where, "the deflation subspace is not initialised (dfl_subspace or dfl_modes)");
      return false;
   }
   return true;
}

}  // namespace

extern "C" void dfl_subspace(spinor **mds)
{
   const char *where = "dfl_subspace [dfl_subspace.c]";
   Lib &L = lib();
   if (!dfl_grid(where)) return;
   HIP_OK(hipMemsetAsync(D.d_unsafe, 0, sizeof(int), L.stream));
   for (int l = 0; l < D.Ns; l++)
   {
      Res r = resolve(mds[l], F32, L.g.V, 0, true);
      if (!r.dev) return;
      if (r.off)
      {
         fail(where, "the modes must be full fields");
         return;
      }
      if (D.bvol <= DWG)
         k_blk_gs_reg<<<D.nb, DWG, 0, L.stream>>>(D.d_basis, D.fu, D.Ns, D.d_bpts, D.bvol, (const float2 *)r.dev, l,
                                                  D.d_vmds, D.d_unsafe);
      else
         k_blk_gs<<<D.nb, DWG, 0, L.stream>>>(D.d_basis, D.fu, D.Ns, D.d_bpts, D.bvol, (const float2 *)r.dev,
                                              D.d_s64, l, D.d_vmds, D.d_unsafe);
      finish(r, F32, L.g.V, false);
   }
   int degen = 0;
   HIP_OK(hipMemcpyAsync(&degen, D.d_unsafe, sizeof(int), hipMemcpyDeviceToHost, L.stream));
   HIP_OK(hipStreamSynchronize(L.stream));
   if (degen)
   {
      fail(where, "Degenerate block modes");
      return;
   }
   D.basis_ok = true;
   D.aw_ok = D.awhat_ok = false;
}

extern "C" void dfl_s2vd(spinor *s, complex_dble *vd)
{
   const char *where = "dfl_s2vd [dfl_subspace.c]";
   Lib &L = lib();
   if (!ready(where)) return;
   Res r = resolve(s, F32, L.g.V, 0, true);
   if (!r.dev) return;
   double2 *w = wvd(0);
   s2vd_dev((const float2 *)r.dev, w);
   HIP_OK(hipMemcpyAsync(vd, w, sizeof(double2) * (size_t)D.nb * D.Ns, hipMemcpyDeviceToHost, L.stream));
   HIP_OK(hipStreamSynchronize(L.stream));
   finish(r, F32, L.g.V, false);
}

extern "C" void dfl_vd2s(complex_dble *vd, spinor *s)
{
   const char *where = "dfl_vd2s [dfl_subspace.c]";
   Lib &L = lib();
   if (!ready(where)) return;
   Res r = resolve(s, F32, L.g.V, 0, false);
   if (!r.dev) return;
   double2 *w = wvd(0);
   HIP_OK(hipMemcpyAsync(w, vd, sizeof(double2) * (size_t)D.nb * D.Ns, hipMemcpyHostToDevice, L.stream));
   vd2s_dev(w, (float2 *)r.dev);
   finish(r, F32, L.g.V, true);
}

// Gram-Schmidt of the global modes (dfl/dfl_subspace.c:261-276).  Device
// fields: per mode one pass forms its products with the previous modes and
// the Gram row of the newest one, a triangular solve turns them into the
// coefficients of the modified process, and one pass applies them and
// normalises (the compact form of the solver's Gram-Schmidt); host arrays
// take the reference's mode-by-mode project/normalize
extern "C" void dfl_renormalize_modes(spinor **mds)
{
   Lib &L = lib();
   const int vt = L.g.VT, Ns = D.Ns;
   std::vector<float2 *> q(Ns, nullptr);
   bool dev = Ns > 0;
   for (int l = 0; l < Ns && dev; l++)
   {
      int off = 0;
      q[l] = (float2 *)device_field(mds[l], F32, &off);
      dev = q[l] != nullptr && off == 0;
   }
   float2 *&coef = D.d_gcoef, *&gram = D.d_ggram, *&scr = D.d_gscr;
   int &cap = D.gcap;
   if (dev && cap < Ns)
   {
      dev_free(coef);
      dev_free(gram);
      dev_free(scr);
      coef = (float2 *)dev_alloc(sizeof(float2) * (Ns + 3));
      gram = (float2 *)dev_alloc(sizeof(float2) * Ns * Ns);
      scr = (float2 *)dev_alloc(sizeof(float2) * units_for(L.g.V, 12));
      cap = Ns;
      if (!coef || !gram || !scr)
         cap = 0;
      else  // the solver's residual slot: zero, so it stays zero
         HIP_OK(hipMemsetAsync(scr, 0, sizeof(float2) * units_for(L.g.V, 12), L.stream));
   }
   if (!dev || cap == 0)
   {
      for (int l = 0; l < Ns; l++)
      {
         for (int j = 0; j < l; j++) project(vt, 3, mds[l], mds[j]);
         (void)normalize(vt, 3, mds[l]);
      }
      return;
   }
   for (int l = 0; l < Ns; l++) gs_classical(q[l], q.data(), l, scr, coef, cap, cap + 1, cap + 2, L.g.V, gram, cap);
}

// basis vector k as a host array in the reference's layout (test access)
// mds[l] = the generating mode l rebuilt from its coordinates in the
// current subspace (dfl/dfl_subspace.c dfl_restore_modes)
extern "C" void dfl_restore_modes(spinor **mds)
{
   const char *where = "dfl_restore_modes [dfl_subspace.c]";
   Lib &L = lib();
   if (!ready(where)) return;
   const size_t nw = (size_t)D.nb * D.Ns;
   double2 *w = wvd(0);
   for (int l = 0; l < D.Ns; l++)
   {
      int off = 0;
      float2 *f = (float2 *)device_field(mds[l], F32, &off);
      if (!f || off)
      {
         fail(where, "the modes must be full device fields");
         return;
      }
      vconv<float2, double2>(D.d_vmds + (size_t)l * nw, w, nw);
      vd2s_dev(w, f);
   }
   (void)L;
}

extern "C" void oqh_dfl_export_basis(int k, spinor *host)
{
   const char *where = "oqh_dfl_export_basis";
   Lib &L = lib();
   if (!ready(where)) return;
   if (k < 0 || k >= D.Ns)
   {
      fail(where, "k out of range");
      return;
   }
   // the basis fields are internal: convert and copy them here
   const size_t bytes = (size_t)L.g.V * 12 * sizeof(float2);
   float2 *aos = (float2 *)tmp_buffer(bytes);
   if (!aos) return;
   blocks_to_aos_f32(D.d_basis + (size_t)k * D.fu, aos, L.g.V, 12, L.stream);
   HIP_OK(hipMemcpyAsync(host, aos, bytes, hipMemcpyDeviceToHost, L.stream));
   HIP_OK(hipStreamSynchronize(L.stream));
}

// block structure: inn[nb][8] and the points of every block, bpts[nb][bvol]
extern "C" int oqh_dfl_blocks(int *inn, int *bpts)
{
   if (!dfl_grid("oqh_dfl_blocks")) return 0;
   if (inn) memcpy(inn, D.inn.data(), sizeof(int) * D.inn.size());
   if (bpts) memcpy(bpts, D.bpts.data(), sizeof(int) * D.bpts.size());
   return D.nb;
}

// ---------------------------------------------------------------------------
// C ABI: the little Dirac operator (little/Aw_ops.c)
// ---------------------------------------------------------------------------

extern "C" void set_Aw(double mu)
{
   const char *where = "set_Aw [Aw_ops.c]";
   Lib &L = lib();
   if (!ready(where)) return;
   if (!L.ud_set)
   {
      fail(where, "gauge field is not set");
      return;
   }
   if (D.aw_ok && D.aw_mu == mu && D.aw_ud == L.ud_version && D.aw_swd == L.swd_epoch && D.aw_eo == L.eoflg &&
       L.swd_valid && !L.swd_inv[0] && !L.swd_inv[1])
      return;
   ensure_swd(0, 0);
   AwArgs a;
   memset(&a, 0, sizeof(a));
   a.basis = D.d_basis;
   a.fu = D.fu;
   a.Ns = D.Ns;
   a.bvol = D.bvol;
   a.V = L.g.V;
   a.h = L.g.V / 2;
   a.masked = (L.bc != 3);
   a.eoflg = L.eoflg;
   a.bpts = D.d_bpts;
   a.nbr = L.d_nbr;
   a.bnd = L.d_bnd;
   a.ud = L.d_ud;
   a.swd = L.d_swd;
   a.mu = mu;
   a.Ablk = D.d_Ablk;
   a.Ahop = D.d_Ahop;
   for (int l = 0; l < D.Ns; l++) k_aw_column<<<D.nb, DWG, 0, L.stream>>>(a, D.d_blk, l);
   HIP_OK(hipStreamSynchronize(L.stream));
   D.aw_ok = true;
   D.awhat_ok = false;
   D.aw_mu = mu;
   D.aw_ud = L.ud_version;
   D.aw_swd = L.swd_epoch;
   D.aw_eo = L.eoflg;
}

extern "C" int set_Awhat(double mu)
{
   Lib &L = lib();
   set_Aw(mu);
   if (!D.aw_ok) return 1;
   if (D.awhat_ok) return 0;
   const size_t lds = sizeof(double2) * (size_t)D.Ns * 2 * D.Ns;
   HIP_OK(hipMemsetAsync(D.d_unsafe, 0, sizeof(int), L.stream));
   k_aw_invert<<<D.nb, 64, lds, L.stream>>>(D.d_Ablk, D.d_Ahop, D.Ns, D.d_AinvT, D.d_MT, D.d_MT32, D.d_unsafe);
   int unsafe = 0;
   HIP_OK(hipMemcpyAsync(&unsafe, D.d_unsafe, sizeof(int), hipMemcpyDeviceToHost, L.stream));
   HIP_OK(hipStreamSynchronize(L.stream));
   if (!unsafe) unsafe = set_ltl_modes();
   D.awhat_ok = (unsafe == 0);
   return unsafe ? 1 : 0;
}

// Ablk [nb][Ns][Ns] and Ahop [nb][8][Ns][Ns] (complex double, row-major) of
// the last set_Aw (test access)
extern "C" void oqh_dfl_export_Aw(complex_dble *ablk, complex_dble *ahop)
{
   Lib &L = lib();
   if (!D.aw_ok)
   {
      fail("oqh_dfl_export_Aw", "the little Dirac operator is not set (set_Aw)");
      return;
   }
   const size_t nm = (size_t)D.Ns * D.Ns;
   if (ablk) HIP_OK(hipMemcpy(ablk, D.d_Ablk, sizeof(double2) * nm * D.nb, hipMemcpyDeviceToHost));
   if (ahop) HIP_OK(hipMemcpy(ahop, D.d_Ahop, sizeof(double2) * nm * 8 * D.nb, hipMemcpyDeviceToHost));
   (void)L;
}

// w = Ahat v on the even blocks (double precision), host arrays of
// nb/2 * Ns complex numbers (little/Aw_dble.c Awhat_dble)
extern "C" void Awhat_dble(complex_dble *v, complex_dble *w)
{
   Lib &L = lib();
   if (!D.awhat_ok)
   {
      fail("Awhat_dble [Aw_dble.c]", "the even-odd little operator is not set (set_Awhat)");
      return;
   }
   const size_t nv = (size_t)D.nbh * D.Ns;
   double2 *a = wvd(1), *t = wvd(2), *b = wvd(3);
   HIP_OK(hipMemcpyAsync(a, v, sizeof(double2) * nv, hipMemcpyHostToDevice, L.stream));
   ahat<double2>(D.d_MT, a, t, b);
   HIP_OK(hipMemcpyAsync(w, b, sizeof(double2) * nv, hipMemcpyDeviceToHost, L.stream));
   HIP_OK(hipStreamSynchronize(L.stream));
}

// ---------------------------------------------------------------------------
// the little solver (dfl/ltl_gcr.c)
// ---------------------------------------------------------------------------

namespace {

// psi = A^-1 eta (full little vectors on the device), FGCR on the even-odd
// form; status[0]: outer Krylov vectors, status[1]: inner ones (sum)
double ltl_solve(const double2 *eta, double2 *psi, int *ifail, int *status)
{
   Lib &L = lib();
   const size_t nv = (size_t)D.nbh * D.Ns;
   ifail[0] = 0;
   status[0] = status[1] = 0;
   const int nkv = D.nkv;
   double2 *u = wvd(4), *b = wvd(5), *x = wvd(6), *r = wvd(7), *t = wvd(8);
   // u_o = Aoo^-1 eta_o, b = Aee^-1 eta_e - sum M_e u_o
   k_aw_diag<<<D.nbh, 64, 0, L.stream>>>(D.d_AinvT, D.Ns, D.nbh, eta + nv, u);
   k_aw_diag<<<D.nbh, 64, 0, L.stream>>>(D.d_AinvT, D.Ns, 0, eta, t);
   aw_hop<double2>(D.d_MT, 0, D.nbh, u, t, b, 1);
   double bn2;
   vdot<double2>(b, b, nv, nullptr, nullptr, &bn2);
   const double rn0 = sqrt(bn2);
   HIP_OK(hipMemsetAsync(x, 0, sizeof(double2) * nv, L.stream));
   // the part along the little modes, x0 = P0 M P0 b, is solved exactly;
   // FGCR then solves P_L Ahat y = P_L b (ltl_gcr.c with ltl_modes.c)
   double2 *x0 = wvd(9);
   double rn = rn0;
   if (D.lm_ok && rn0 > 0.0)
   {
      dfl_LRvd_dev(b, x0);
      vdot<double2>(b, b, nv, nullptr, nullptr, &bn2);
      rn = sqrt(bn2);
   }
   if (rn0 > 0.0)
   {
      // FGCR (linsolv/fgcr4vd.c) with Mop = gcr4v in single precision
      if (!lkry_alloc(K64, nkv, nv))
      {
         ifail[0] = -1;
         return rn;
      }
      float2 *b32 = wv(2), *x32 = wv(3);
      HIP_OK(hipMemcpyAsync(r, b, sizeof(double2) * nv, hipMemcpyDeviceToDevice, L.stream));
      const double tol = D.res * rn0;
      while (rn > tol)
      {
         int k;
         bool zero = false;
         for (k = 0;; k++)
         {
            // phi = (approximate Ahat^-1) r, chi = Ahat phi
            vconv<double2, float2>(r, b32, nv);
            const double rres = (rn * D.res_gcr < tol) ? tol / rn : D.res_gcr;
            const int st = gcr4v(D.nmx_gcr, rres, b32, x32);
            status[1] += (st < 0) ? D.nmx_gcr : st;
            vconv<float2, double2>(x32, K64.phi(k), nv);
            ahat<double2>(D.d_MT, K64.phi(k), t, K64.chi(k));
            if (D.lm_ok) dfl_RLvd_dev(K64.phi(k), K64.chi(k));
            if (!lgs_step(K64, k, r, &rn))
            {
               zero = true;
               break;
            }
            status[0] += 1;
            if (rn <= tol || k == nkv - 1 || status[0] >= D.nmx) break;
         }
         lgs_solution(K64, zero ? k : k + 1, x, 1);
         // recompute the residual r = P_L (b - Ahat x)
         ahat<double2>(D.d_MT, x, t, r);
         if (D.lm_ok) dfl_Lvd_dev(r);
         k_vscale<double2><<<(unsigned)((nv + 255) / 256), 256, 0, L.stream>>>(r, nv, -1.0);
         vaxpy<double2>(r, b, nv, 1.0, 0.0);
         double r2;
         vdot<double2>(r, r, nv, nullptr, nullptr, &r2);
         rn = sqrt(r2);
         if (status[0] >= D.nmx && rn > tol)
         {
            ifail[0] = -1;
            break;
         }
         if (zero) break;
      }
   }
   if (D.lm_ok && rn0 > 0.0) vaxpy<double2>(x, x0, nv, 1.0, 0.0);
   // psi_e = x, psi_o = u_o - sum M_o x_e
   HIP_OK(hipMemcpyAsync(psi, x, sizeof(double2) * nv, hipMemcpyDeviceToDevice, L.stream));
   aw_hop<double2>(D.d_MT, D.nbh, 0, x, u, psi + nv, 1);
   return rn;
}

}  // namespace

extern "C" double ltl_gcr(double mu, complex_dble *eta, complex_dble *psi, int *ifail, int *status)
{
   const char *where = "ltl_gcr [ltl_gcr.c]";
   Lib &L = lib();
   if (!ready(where)) return 0.0;
   if (D.nkv == 0)
   {
      fail(where, "Deflation projection parameters are not set");
      return 0.0;
   }
   const size_t n = (size_t)D.nb * D.Ns;
   if (set_Awhat(mu))
   {
      ifail[0] = -2;
      status[0] = status[1] = 0;
      return 0.0;
   }
   double2 *e = wvd(10), *p = wvd(11);
   HIP_OK(hipMemcpyAsync(e, eta, sizeof(double2) * n, hipMemcpyHostToDevice, L.stream));
   const double rho = ltl_solve(e, p, ifail, status);
   HIP_OK(hipMemcpyAsync(psi, p, sizeof(double2) * n, hipMemcpyDeviceToHost, L.stream));
   HIP_OK(hipStreamSynchronize(L.stream));
   return rho;
}

// ---------------------------------------------------------------------------
// the deflated solver (dfl/dfl_sap_gcr.c) and the generation of the
// subspace (dfl/dfl_modes.c)
// ---------------------------------------------------------------------------

namespace oqh {
void random_uniform_s(float2 *f);  // rng.hip: uniform in [-1/2, 1/2) per component
}

namespace {

struct DflRun
{
   double mud = 0.0;
   float mus = 0.0f;
   int isolv = 1, nmr = 4, ncy = 5;
   int ltl[2] = {0, 0};
   float2 *psi0 = nullptr;   // FP32 work field of the preconditioner
   float2 *wtmp = nullptr;
   double2 *vin = nullptr, *vout = nullptr;
};

DflRun R;

// a field of the library's registry, so that the operators and the BLAS of
// the C API accept it as a device field
float2 *work_field(float2 **p)
{
   if (!*p) *p = (float2 *)oqh_alloc_s(lib().g.V);
   return *p;
}

// the deflated preconditioner (dfl_sap_gcr.c Mop): psi0 = the solution of
// the little equation for the components of rho along the subspace, then ncy
// Schwarz cycles on the residual r1 = rho - D psi0; phi = psi0 + SAP(r1),
// chi = D phi.  r1 lives in a fixed work field, so that the FP16 variant's
// captured Schwarz cycles are replayed for every Krylov vector.
void dfl_mop_dev(const float2 *rho, float2 *phi, float2 *chi)
{
   Lib &L = lib();
   const int vt = L.g.VT;
   float2 *p0 = work_field(&R.psi0), *r1 = work_field(&R.wtmp);
   if (!R.vin) R.vin = wvd(12);
   if (!R.vout) R.vout = wvd(13);
   s2vd_dev(rho, R.vin);
   int ifail = 0, st[2] = {0, 0};
   (void)ltl_solve(R.vin, R.vout, &ifail, st);
   R.ltl[0] += st[0];
   R.ltl[1] += st[1];
   vd2s_dev(R.vout, p0);
   Dw(R.mus, (spinor *)p0, (spinor *)r1);
   diff_s2s(vt, 2, (spinor *)rho, (spinor *)r1);   // r1 = rho - D psi0
   if (L.sap_prec == 1 && sap_mixed_mop(R.mus, R.isolv, R.nmr, R.ncy, r1, phi, false))
   {
      // FP16 Schwarz cycles from psi = 0; chi = D (phi + psi0) in FP32
      add_s2s(vt, 2, (spinor *)p0, (spinor *)phi);
      Dw(R.mus, (spinor *)phi, (spinor *)chi);
      return;
   }
   set_s2zero(vt, 2, (spinor *)phi);
   for (int n = 0; n < R.ncy; n++) sap_device(R.mus, R.isolv, R.nmr, phi, r1);
   // r1 = rho - D psi0 - D phi_sap, so chi = rho - r1 = D (psi0 + phi_sap)
   assign_s2s(vt, 2, (spinor *)r1, (spinor *)chi);
   diff_s2s(vt, 2, (spinor *)rho, (spinor *)chi);
   add_s2s(vt, 2, (spinor *)p0, (spinor *)phi);
}

void dfl_Dop(spinor_dble *s, spinor_dble *r) { Dw_dble(R.mud, s, r); }

void dfl_Mop(int k, spinor *rho, spinor *phi, spinor *chi)
{
   (void)k;
   int o0 = 0, o1 = 0, o2 = 0;
   float2 *r = (float2 *)device_field(rho, F32, &o0), *p = (float2 *)device_field(phi, F32, &o1),
          *c = (float2 *)device_field(chi, F32, &o2);
   if (!r || !p || !c || o0 || o1 || o2)
   {
      fail("dfl_sap_gcr [dfl_sap_gcr.c]", "the Krylov vectors must be full device fields");
      return;
   }
   dfl_mop_dev(r, p, c);
}

}  // namespace

extern "C" double dfl_sap_gcr(int nkv, int nmx, int istop, double res, double mu, spinor_dble *eta,
                              spinor_dble *psi, int *ifail, int *status)
{
   const char *where = "dfl_sap_gcr [dfl_sap_gcr.c]";
   Lib &L = lib();
   ifail[0] = 0;
   status[0] = status[1] = status[2] = 0;
   if (!ready(where)) return 0.0;
   if (L.sap_ncy == 0)
   {
      fail(where, "SAP parameters are not set");
      return 0.0;
   }
   if (D.nkv == 0)
   {
      fail(where, "Deflation projection parameters are not set");
      return 0.0;
   }
   const int vt = L.g.VT;
   double rho = 0.0;
   if (!sap_prepare(L.sap_isolv, where, true))
      ifail[0] = -3;
   else if (set_Awhat(mu))
      ifail[0] = -2;
   else
   {
      const double rho0 = istop ? unorm_dble(vt, 3, eta) : sqrt(norm_square_dble(vt, 3, eta).q[0]);
      if (rho0 != 0.0)
      {
         alloc_ws(L.ws_used + 2 * nkv + 1);
         alloc_wsd(L.wsd_used + 2);
         spinor **ws = reserve_ws(2 * nkv + 1);
         spinor_dble **wsd = reserve_wsd(2);
         spinor_dble *rsd = wsd[1];
         assign_sd2sd(vt, 2, eta, rsd);
         scale_dble(vt, 2, 1.0 / rho0, rsd);
         R.mud = mu;
         R.mus = (float)mu;
         R.isolv = L.sap_isolv;
         R.nmr = L.sap_nmr;
         R.ncy = L.sap_ncy;
         R.ltl[0] = R.ltl[1] = 0;
         rho = fgcr(vt, dfl_Dop, dfl_Mop, ws, wsd, nkv, nmx, istop, res, rsd, psi, status);
         scale_dble(vt, 2, rho0, psi);
         rho *= rho0;
         if (status[0] < 0)
         {
            ifail[0] = -1;
            status[0] = nmx;
         }
         if (status[0] > 0)
         {
            status[1] = (R.ltl[0] + status[0] / 2) / status[0];
            status[2] = (R.ltl[1] + status[0] / 2) / status[0];
         }
         release_wsd();
         release_ws();
      }
      else
         set_sd2zero(vt, 2, psi);
   }
   if (ifail[0] <= -2)
   {
      if (psi != eta) set_sd2zero(vt, 2, psi);
      rho = istop ? unorm_dble(vt, 3, eta) : sqrt(norm_square_dble(vt, 3, eta).q[0]);
   }
   return rho;
}

namespace {

// checks and set_flds of dfl_modes.c: the generation mass and the
// twisted-mass flag (restored by gen_clean_up); false if a check failed
struct GenState
{
   double m0_keep = 0.0;
   int eo_keep = 0;
};

bool gen_prepare(const char *where, GenState &g, int *ifail)
{
   Lib &L = lib();
   if (!dfl_grid(where)) return false;
   if (D.nkv == 0 || D.ninv == 0)
   {
      fail(where, "Deflation projection or subspace generation parameters are not set");
      return false;
   }
   if (L.sap_ncy == 0)
   {
      fail(where, "SAP parameters are not set");
      return false;
   }
   if (!L.ud_set)
   {
      fail(where, "gauge field is not set");
      return false;
   }
   g.m0_keep = L.m0;
   g.eo_keep = L.eoflg;
   if (D.m0 != L.m0) oqh_set_sw_parms(D.m0, L.csw);
   oqh_set_tm_parms(1);
   if (!sap_prepare(1, where, true)) ifail[0] = -3;
   return true;
}

void gen_clean_up(const GenState &g)
{
   Lib &L = lib();
   if (L.m0 != g.m0_keep) oqh_set_sw_parms(g.m0_keep, L.csw);
   oqh_set_tm_parms(g.eo_keep);
   D.aw_ok = D.awhat_ok = false;
}

// ncy Schwarz cycles from psi = 0 on rho = r (the smoother of the subspace
// generation): with oqh_set_sap_prec(1) on FP16 fields (HIP graph keyed on
// the fixed field r), else the FP32 cycles of sap.c
void gen_cycles(float mus, int ncy, spinor *psi, spinor *r)
{
   Lib &L = lib();
   int off = 0;
   float2 *p = (float2 *)device_field(psi, F32, &off), *rr = (float2 *)device_field(r, F32, &off);
   if (L.sap_prec == 1 && sap_mixed_mop(mus, 1, D.nmr, ncy, rr, p, false)) return;
   set_s2zero(L.g.VT, 2, psi);
   for (int j = 0; j < ncy; j++) sap_device(mus, 1, D.nmr, p, rr);
}

// one deflated inverse-iteration step of every mode (dfl_smooth_fields):
// mds_l <- psi1 + ncy Schwarz cycles on mds_l - D psi1, psi1 the prolongated
// little solution of the mode's coordinates; adds the little solver's
// Krylov counts to st
void dfl_smooth_fields(spinor **mds, spinor **ws, int *st)
{
   Lib &L = lib();
   const int vt = L.g.VT;
   const float mus = (float)D.mu;
   const size_t nw = (size_t)D.nb * D.Ns;
   double2 *w = wvd(14), *w2 = wvd(15);
   int off = 0;
   for (int l = 0; l < D.Ns; l++)
   {
      vconv<float2, double2>(D.d_vmds + (size_t)l * nw, w, nw);
      int lf = 0, s2[2] = {0, 0};
      (void)ltl_solve(w, w2, &lf, s2);
      st[0] += s2[0];
      st[1] += s2[1];
      float2 *p1 = (float2 *)device_field(ws[1], F32, &off);
      vd2s_dev(w2, p1);
      Dw(mus, ws[1], ws[0]);
      diff_s2s(vt, 2, mds[l], ws[0]);   // ws0 = mds - D psi1
      gen_cycles(mus, D.ncy, mds[l], ws[0]);
      add_s2s(vt, 2, ws[1], mds[l]);
   }
}

// n deflated steps (set_Awhat, dfl_smooth_fields, renormalisation, new
// subspace) while ifail[0] == 0; returns the number of steps done
int dfl_steps(int n, spinor **mds, spinor **ws, int *ifail, int *status)
{
   int k = 0;
   for (; k < n && ifail[0] == 0; k++)
   {
      if (set_Awhat(D.mu))
      {
         ifail[0] = -2;
         break;
      }
      dfl_smooth_fields(mds, ws, status);
      dfl_renormalize_modes(mds);
      dfl_subspace(mds);
   }
   return k;
}

void average_status(int nsolve, int *status)
{
   if (nsolve > 0)
   {
      status[0] = (status[0] + nsolve / 2) / nsolve;
      status[1] = (status[1] + nsolve / 2) / nsolve;
   }
}

bool gen_done = false;   // a subspace was generated by dfl_modes

}  // namespace

extern "C" void dfl_modes(int *ifail, int *status)
{
   const char *where = "dfl_modes [dfl_modes.c]";
   Lib &L = lib();
   ifail[0] = 0;
   status[0] = status[1] = 0;
   GenState g;
   if (!gen_prepare(where, g, ifail)) return;
   const int vt = L.g.VT;
   alloc_ws(L.ws_used + D.Ns + 2);
   spinor **mds = reserve_ws(D.Ns);
   spinor **ws = reserve_ws(2);
   int off = 0;  // site offset of a device field (the fields are full)
   for (int l = 0; l < D.Ns; l++)
   {
      random_uniform_s((float2 *)device_field(mds[l], F32, &off));
      bnd_s2zero(ALL_PTS, mds[l]);
   }
   const float mus = (float)D.mu;
   if (ifail[0] != 0)
      dfl_subspace(mds);
   else
   {
      // smooth_fields: n+1 Schwarz cycles on D psi = mds[l], n = 0, 1, 2
      for (int n = 0; n < 3; n++)
         for (int l = 0; l < D.Ns; l++)
         {
            assign_s2s(vt, 2, mds[l], ws[0]);
            gen_cycles(mus, n + 1, mds[l], ws[0]);
         }
      dfl_subspace(mds);
      const int nd = dfl_steps(D.ninv - 3, mds, ws, ifail, status);
      average_status(D.Ns * nd, status);
      gen_done = true;
   }
   release_ws();
   release_ws();
   gen_clean_up(g);
}

// nsm further deflated inverse-iteration steps on the modes restored from
// the current subspace (dfl/dfl_modes.c dfl_update)
extern "C" void dfl_update(int nsm, int *ifail, int *status)
{
   const char *where = "dfl_update [dfl_modes.c]";
   Lib &L = lib();
   ifail[0] = 0;
   status[0] = status[1] = 0;
   if (!gen_done || !D.basis_ok)
   {
      fail(where, "No previous call of dfl_modes()");
      return;
   }
   GenState g;
   if (!gen_prepare(where, g, ifail)) return;
   if (ifail[0] == 0)
   {
      alloc_ws(L.ws_used + D.Ns + 2);
      spinor **mds = reserve_ws(D.Ns);
      spinor **ws = reserve_ws(2);
      dfl_restore_modes(mds);
      const int nd = dfl_steps(nsm, mds, ws, ifail, status);
      average_status(D.Ns * nd, status);
      release_ws();
      release_ws();
   }
   gen_clean_up(g);
}

extern "C" void dfl_update2(int nsm, int *ifail, int *status)
{
   dfl_update(nsm, ifail, status);
   if (ifail[0] == -2)
      dfl_modes(ifail + 1, status + 2);
   else
   {
      ifail[1] = 0;
      status[2] = status[3] = 0;
   }
}

extern "C" void dfl_modes2(int *ifail, int *status)
{
   dfl_modes(ifail, status);
   if (ifail[0] == -2)
      dfl_modes(ifail + 1, status + 2);
   else
   {
      ifail[1] = 0;
      status[2] = status[3] = 0;
   }
}

extern "C" double dfl_sap_gcr2(int nkv, int nmx, int istop, double res, double mu, spinor_dble *eta,
                               spinor_dble *psi, int *ifail, int *status)
{
   Lib &L = lib();
   const int vt = L.g.VT;
   ifail[1] = 0;
   for (int l = 3; l < 6; l++) status[l] = 0;
   alloc_wsd(L.wsd_used + 1);
   spinor_dble *rsd = reserve_wsd(1)[0];
   assign_sd2sd(vt, 2, eta, rsd);
   double rho = dfl_sap_gcr(nkv, nmx, istop, res, mu, rsd, psi, ifail, status);
   if (ifail[0] == -1 || ifail[0] == -2)
   {
      dfl_modes(ifail + 1, status + 4);
      if (ifail[1] == 0) rho = dfl_sap_gcr(nkv, nmx, istop, res, mu, rsd, psi, ifail + 1, status + 3);
   }
   release_wsd();
   return rho;
}

// the whole deflation state, parameters included (oqh_finalize)
void oqh::release_dfl()
{
   dfl_free_grid();
   if (R.psi0) oqh_free(R.psi0);
   if (R.wtmp) oqh_free(R.wtmp);
   R = DflRun();
   D = Dfl();
}

extern "C" void oqh_dfl_release(void) { release_dfl(); }
