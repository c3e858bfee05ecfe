// linalg.hip -- spinor BLAS-1, field assignments, uniform norms and the
// double-double sums, on gfx950.
//
// Replaces linalg/salg.c:1223-1429, linalg/salg_dble.c:2611-2891,
// sflds/sflds.c:1203-1405, sflds/unorm.c:261-314 and utils/qsum.c:262-424.
//
// Reductions are deterministic: every work-item sums a fixed set of points in
// double precision (grid-stride order), the partial sums are combined as
// double-double numbers (Knuth's two-sum, as acc_qflt/add_qflt do) first
// across the wavefront with lane shuffles, then across the work-group through
// LDS, and finally over the work-group partials in a fixed order by a single
// work-group.  The FP64 products therefore end in a (hi,lo) pair whose sum is
// at least as accurate as the reference's chunked accumulation.  The FP32
// products follow salg.c: per point the colour-vector products are formed in
// float, converted to double and summed in double; the result is rounded to
// float.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "oqh_internal.h"
#include "ddsum.h"

namespace oqh {

constexpr int RB = 256;      // reduction work-group size
constexpr int RMAXB = 1024;  // maximum number of reduction work-groups

// per-point contribution of the supported products
enum DotKind { DOT_C = 0, DOT_RE, DOT_NORM, DOT_C5 };

template <typename C, int KIND>
__device__ inline void point_dot(const C *a, const C *b, int x, int y, double *re, double *im)
{
   const C *pa = a + (size_t)(x >> 6) * (12 * SB) + (x & (SB - 1));
   const C *pb = b + (size_t)(y >> 6) * (12 * SB) + (y & (SB - 1));
   typedef decltype(C::x) R;
   if (sizeof(R) == 8)
   {
      double sr = 0.0, si = 0.0;
#pragma unroll
      for (int c = 0; c < 12; c++)
      {
         C u = pa[c * SB], v = (KIND == DOT_NORM) ? u : pb[c * SB];
         double g = (KIND == DOT_C5 && c >= 6) ? -1.0 : 1.0;
         sr += g * ((double)u.x * v.x + (double)u.y * v.y);
         if (KIND == DOT_C || KIND == DOT_C5) si += g * ((double)u.x * v.y - (double)u.y * v.x);
      }
      *re += sr;
      *im += si;
   }
   else
   {
      // salg.c: per colour vector in float, then in double
#pragma unroll
      for (int k = 0; k < 4; k++)
      {
         R pr = 0.0f, pi = 0.0f;
#pragma unroll
         for (int i = 0; i < 3; i++)
         {
            C u = pa[(3 * k + i) * SB], v = (KIND == DOT_NORM) ? u : pb[(3 * k + i) * SB];
            pr += u.x * v.x + u.y * v.y;
            if (KIND == DOT_C) pi += u.x * v.y - u.y * v.x;
         }
         *re += (double)pr;
         *im += (double)pi;
      }
   }
}

template <typename C, int KIND>
__global__ __launch_bounds__(RB) void k_dot_partial(const C *a, const C *b, int aoff, int boff, int n, dd *part)
{
   double re = 0.0, im = 0.0;
   int stride = gridDim.x * RB;
   for (int k = blockIdx.x * RB + threadIdx.x; k < n; k += stride)
      point_dot<C, KIND>(a, b, aoff + k, boff + k, &re, &im);
   dd v[2] = {{re, 0.0}, {im, 0.0}};
   block_reduce_dd<2>(v);
   if (threadIdx.x == 0)
   {
      part[2 * blockIdx.x] = v[0];
      part[2 * blockIdx.x + 1] = v[1];
   }
}

__global__ __launch_bounds__(RB) void k_dd_final(const dd *part, int nb, dd *out)
{
   dd v[2] = {{0.0, 0.0}, {0.0, 0.0}};
   for (int k = threadIdx.x; k < nb; k += RB)
   {
      v[0] = dd_add(v[0], part[2 * k]);
      v[1] = dd_add(v[1], part[2 * k + 1]);
   }
   block_reduce_dd<2>(v);
   if (threadIdx.x == 0)
   {
      out[0] = v[0];
      out[1] = v[1];
   }
}

static dd *d_part = nullptr, *h_res = nullptr;

static int red_blocks(int n)
{
   int nb = (n + RB - 1) / RB;
   return nb < 1 ? 1 : (nb > RMAXB ? RMAXB : nb);
}

// launches the reduction; result (re, im) as double-double in res[0..1]
template <typename C, int KIND>
void dot_async(const C *a, const C *b, int aoff, int boff, int n, dd *d_out, hipStream_t st)
{
   if (!d_part) d_part = (dd *)dev_alloc(sizeof(dd) * 2 * (RMAXB + 8));
   int nb = red_blocks(n);
   k_dot_partial<C, KIND><<<nb, RB, 0, st>>>(a, b, aoff, boff, n, d_part);
   k_dd_final<<<1, RB, 0, st>>>(d_part, nb, d_out);
}

static dd *d_out_scalar = nullptr;

template <typename C, int KIND> static void dot_sync(const C *a, const C *b, int aoff, int boff, int n, dd *res)
{
   Lib &L = lib();
   if (!d_out_scalar) d_out_scalar = (dd *)dev_alloc(sizeof(dd) * 2);
   if (!h_res) HIP_OK(hipHostMalloc((void **)&h_res, sizeof(dd) * 2, 0));
   dot_async<C, KIND>(a, b, aoff, boff, n, d_out_scalar, L.stream);
   HIP_OK(hipMemcpyAsync(h_res, d_out_scalar, sizeof(dd) * 2, hipMemcpyDeviceToHost, L.stream));
   HIP_OK(hipStreamSynchronize(L.stream));
   res[0] = h_res[0];
   res[1] = h_res[1];
}

// ---------------------------------------------------------------------------
// element-wise kernels (one work-item per point)
// ---------------------------------------------------------------------------

enum AxKind { AX_MULC = 0, AX_MULR, AX_SCALE, AX_G5, AX_MG5, AX_COMB, AX_ZERO, AX_COPY, AX_ADD, AX_DIFF };

template <typename C, int K>
__global__ __launch_bounds__(256) void k_axpy(C *s, const C *r, int soff, int roff, int n,
                                              decltype(C::x) zr, decltype(C::x) zi, decltype(C::x) cr)
{
   int k = blockIdx.x * 256 + threadIdx.x;
   if (k >= n) return;
   int xs = soff + k, xr = roff + k;
   C *ps = s + (size_t)(xs >> 6) * (12 * SB) + (xs & (SB - 1));
   const C *pr = r ? r + (size_t)(xr >> 6) * (12 * SB) + (xr & (SB - 1)) : nullptr;
#pragma unroll
   for (int c = 0; c < 12; c++)
   {
      C a = ps[c * SB];
      if (K == AX_MULC)
      {
         C b = pr[c * SB];
         a = cmk<C>(a.x + (zr * b.x - zi * b.y), a.y + (zr * b.y + zi * b.x));
      }
      else if (K == AX_MULR)
      {
         C b = pr[c * SB];
         a = cmk<C>(a.x + zr * b.x, a.y + zr * b.y);
      }
      else if (K == AX_SCALE)
         a = cmk<C>(zr * a.x, zr * a.y);
      else if (K == AX_G5)
      {
         if (c >= 6) a = cmk<C>(-a.x, -a.y);
      }
      else if (K == AX_MG5)
      {
         if (c < 6) a = cmk<C>(-a.x, -a.y);
      }
      else if (K == AX_COMB)
      {
         C b = pr[c * SB];
         a = cmk<C>(zr * a.x + cr * b.x, zr * a.y + cr * b.y);
      }
      else if (K == AX_ZERO)
         a = cmk<C>(0, 0);
      else if (K == AX_COPY)
         a = pr[c * SB];
      else if (K == AX_ADD)
      {
         C b = pr[c * SB];
         a = cmk<C>(a.x + b.x, a.y + b.y);
      }
      else if (K == AX_DIFF)  // s <- r - s   (diff_s2s: r = s - r with roles swapped by caller)
      {
         C b = pr[c * SB];
         a = cmk<C>(b.x - a.x, b.y - a.y);
      }
      ps[c * SB] = a;
   }
}

// mixed-precision assignments
template <int K>
__global__ __launch_bounds__(256) void k_mixed(const void *a, const void *b, void *r, int aoff, int boff,
                                               int roff, int n)
{
   int k = blockIdx.x * 256 + threadIdx.x;
   if (k >= n) return;
   int xa = aoff + k, xb = boff + k, xr = roff + k;
   size_t ia = (size_t)(xa >> 6) * (12 * SB) + (xa & (SB - 1));
   size_t ib = (size_t)(xb >> 6) * (12 * SB) + (xb & (SB - 1));
   size_t ir = (size_t)(xr >> 6) * (12 * SB) + (xr & (SB - 1));
#pragma unroll
   for (int c = 0; c < 12; c++)
   {
      if (K == 0)  // assign_s2sd: float a -> double r
      {
         float2 v = ((const float2 *)a)[ia + c * SB];
         ((double2 *)r)[ir + c * SB] = make_double2(v.x, v.y);
      }
      else if (K == 1)  // assign_sd2s: double a -> float r
      {
         double2 v = ((const double2 *)a)[ia + c * SB];
         ((float2 *)r)[ir + c * SB] = make_float2((float)v.x, (float)v.y);
      }
      else if (K == 2)  // add_s2sd: double r += float a
      {
         float2 v = ((const float2 *)a)[ia + c * SB];
         double2 w = ((double2 *)r)[ir + c * SB];
         ((double2 *)r)[ir + c * SB] = make_double2(w.x + (double)v.x, w.y + (double)v.y);
      }
      else  // diff_sd2s: float r = double a - double b
      {
         double2 v = ((const double2 *)a)[ia + c * SB];
         double2 w = ((const double2 *)b)[ib + c * SB];
         ((float2 *)r)[ir + c * SB] = make_float2((float)(v.x - w.x), (float)(v.y - w.y));
      }
   }
}

// per-point squared norm maximum
template <typename C>
__global__ __launch_bounds__(RB) void k_unorm_partial(const C *a, int off, int n, double *part)
{
   __shared__ double sh[RB / 64];
   double m = 0.0;
   int stride = gridDim.x * RB;
   typedef decltype(C::x) R;
   for (int k = blockIdx.x * RB + threadIdx.x; k < n; k += stride)
   {
      int x = off + k;
      const C *p = a + (size_t)(x >> 6) * (12 * SB) + (x & (SB - 1));
      R s = 0;
#pragma unroll
      for (int c = 0; c < 12; c++)
      {
         C u = p[c * SB];
         s += u.x * u.x + u.y * u.y;
      }
      m = fmax(m, (double)s);
   }
#pragma unroll
   for (int sft = 32; sft >= 1; sft >>= 1) m = fmax(m, __shfl_xor(m, sft));
   if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
   __syncthreads();
   if (threadIdx.x == 0)
   {
      for (int w = 1; w < RB / 64; w++) m = fmax(m, sh[w]);
      part[blockIdx.x] = m;
   }
}

__global__ void k_max_final(const double *part, int nb, double *out)
{
   if (threadIdx.x != 0) return;
   double m = 0.0;
   for (int k = 0; k < nb; k++) m = fmax(m, part[k]);
   out[0] = m;
}

// ---------------------------------------------------------------------------
// fused Gram-Schmidt steps of fgcr (single precision, device coefficients)
//
// Each kernel applies one vector update and forms the partial sums of the
// next scalar product in the same pass, with the point partition of
// k_dot_partial (grid-stride over the points, RB-lane work-groups, the same
// number of work-groups), so the products are bitwise those of the separate
// salg.c calls.  The coefficients stay on the device: the update of a kernel
// reads the value the previous reduction stored, and the host reads them back
// once per Krylov vector.
// ---------------------------------------------------------------------------

// conj(u).v per point as salg.c (float per colour vector, then double)
__device__ inline void pdot_reg(const float2 *u, const float2 *v, bool norm, double *re, double *im)
{
#pragma unroll
   for (int k = 0; k < 4; k++)
   {
      float pr = 0.0f, pi = 0.0f;
#pragma unroll
      for (int i = 0; i < 3; i++)
      {
         float2 a = u[3 * k + i], b = norm ? a : v[3 * k + i];
         pr += a.x * b.x + a.y * b.y;
         if (!norm) pi += a.x * b.y - a.y * b.x;
      }
      *re += (double)pr;
      *im += (double)pi;
   }
}

// s, r, q: fields of n points from site 0.  GS_DOT: (q, s).  GS_AXPY_DOT:
// s -= coef[ic] r, then (q, s).  GS_AXPY_NORM: s -= coef[ic] r, then |s|^2.
// GS_SCALE_DOT: s *= 1/coef[ic].x (the norm of s), then (s, q).  GS_NORM: |s|^2.
template <int MODE>
__global__ __launch_bounds__(RB) void k_gs_partial(float2 *s, const float2 *r, const float2 *q, const float2 *coef,
                                                   int ic, int n, dd *part)
{
   double re = 0.0, im = 0.0;
   float zr = 0.0f, zi = 0.0f, sc = 0.0f;
   if (MODE == GS_AXPY_DOT || MODE == GS_AXPY_NORM)
   {
      zr = -coef[ic].x;
      zi = -coef[ic].y;
   }
   if (MODE == GS_SCALE_DOT) sc = 1.0f / coef[ic].x;
   int stride = gridDim.x * RB;
   for (int k = blockIdx.x * RB + threadIdx.x; k < n; k += stride)
   {
      size_t i0 = (size_t)(k >> 6) * (12 * SB) + (k & (SB - 1));
      float2 a[12], b[12];
#pragma unroll
      for (int c = 0; c < 12; c++) a[c] = s[i0 + c * SB];
      if (MODE == GS_AXPY_DOT || MODE == GS_AXPY_NORM)
      {
#pragma unroll
         for (int c = 0; c < 12; c++)
         {
            float2 w = r[i0 + c * SB];
            a[c] = make_float2(a[c].x + (zr * w.x - zi * w.y), a[c].y + (zr * w.y + zi * w.x));
            s[i0 + c * SB] = a[c];
         }
      }
      if (MODE == GS_SCALE_DOT)
      {
#pragma unroll
         for (int c = 0; c < 12; c++)
         {
            a[c] = make_float2(sc * a[c].x, sc * a[c].y);
            s[i0 + c * SB] = a[c];
         }
      }
      if (MODE == GS_DOT || MODE == GS_AXPY_DOT || MODE == GS_SCALE_DOT)
      {
#pragma unroll
         for (int c = 0; c < 12; c++) b[c] = q[i0 + c * SB];
      }
      if (MODE == GS_DOT || MODE == GS_AXPY_DOT)
         pdot_reg(b, a, false, &re, &im);
      else if (MODE == GS_SCALE_DOT)
         pdot_reg(a, b, false, &re, &im);
      else
         pdot_reg(a, a, true, &re, &im);
   }
   dd v[2] = {{re, 0.0}, {im, 0.0}};
   block_reduce_dd<2>(v);
   if (threadIdx.x == 0)
   {
      part[2 * blockIdx.x] = v[0];
      part[2 * blockIdx.x + 1] = v[1];
   }
}

// final sum; stores the float result in coef[io]: the complex product, or
// for a norm (NORM=1) the value sqrt(float(|s|^2)) of normalize() in .x and
// float(|s|^2) in .y
template <int NORM>
__global__ __launch_bounds__(RB) void k_gs_final(const dd *part, int nb, float2 *coef, int io)
{
   dd v[2] = {{0.0, 0.0}, {0.0, 0.0}};
   for (int k = threadIdx.x; k < nb; k += RB)
   {
      v[0] = dd_add(v[0], part[2 * k]);
      v[1] = dd_add(v[1], part[2 * k + 1]);
   }
   block_reduce_dd<2>(v);
   if (threadIdx.x == 0)
   {
      if (NORM)
      {
         float r = (float)(v[0].hi + v[0].lo);
         coef[io] = make_float2((float)sqrt((double)r), r);
      }
      else
         coef[io] = make_float2((float)(v[0].hi + v[0].lo), (float)(v[1].hi + v[1].lo));
   }
}

template <int MODE>
void gs_step(float2 *s, const float2 *r, const float2 *q, float2 *coef, int ic, int io, int n)
{
   Lib &L = lib();
   if (!d_part) d_part = (dd *)dev_alloc(sizeof(dd) * 2 * (RMAXB + 8));
   int nb = red_blocks(n);
   k_gs_partial<MODE><<<nb, RB, 0, L.stream>>>(s, r, q, coef, ic, n, d_part);
   if (MODE == GS_AXPY_NORM || MODE == GS_NORM)
      k_gs_final<1><<<1, RB, 0, L.stream>>>(d_part, nb, coef, io);
   else
      k_gs_final<0><<<1, RB, 0, L.stream>>>(d_part, nb, coef, io);
}

template void gs_step<GS_DOT>(float2 *, const float2 *, const float2 *, float2 *, int, int, int);
template void gs_step<GS_AXPY_DOT>(float2 *, const float2 *, const float2 *, float2 *, int, int, int);
template void gs_step<GS_AXPY_NORM>(float2 *, const float2 *, const float2 *, float2 *, int, int, int);
template void gs_step<GS_SCALE_DOT>(float2 *, const float2 *, const float2 *, float2 *, int, int, int);
template void gs_step<GS_NORM>(float2 *, const float2 *, const float2 *, float2 *, int, int, int);

// ---------------------------------------------------------------------------
// Gram-Schmidt step of fgcr with two passes over the Krylov vectors (the
// inverse compact WY form of modified Gram-Schmidt, Swirydowicz et al.,
// Numer. Linear Algebra Appl. 28 (2021) e2343).  The coefficients of the
// modified process of fgcr.c:157-178,
//    d_i = (chi_i, s - sum_{l<i} d_l chi_l) = p_i - sum_{l<i} G_il d_l,
// with p_i = (chi_i, s) and G_il = (chi_i, chi_l), follow from one pass
// that forms all p_i together with the Gram entries of the newest Krylov
// vector, and a triangular solve of size k; one more pass applies all
// updates.  Unlike classical Gram-Schmidt the non-orthogonality of the
// stored FP32 vectors is corrected as in the modified process.
// ---------------------------------------------------------------------------

constexpr int NQD = 12;  // vectors per product pass
constexpr int NQA = 32;  // vectors per update pass

struct QList
{
   const float2 *q[NQA];
};

static dd *d_mpart = nullptr;

// partial sums of (q_j, s) and, if c != null, (q_j, c), j < nq <= NQD;
// part[blockIdx][4 j + (0,1: with s; 2,3: with c)]
__global__ __launch_bounds__(RB) void k_mdot_partial(const float2 *__restrict__ s, const float2 *__restrict__ c,
                                                     QList ql, int nq, int n, dd *part)
{
   double re[NQD], im[NQD], gr[NQD], gi[NQD];
#pragma unroll
   for (int j = 0; j < NQD; j++) re[j] = im[j] = gr[j] = gi[j] = 0.0;
   int stride = gridDim.x * RB;
   for (int k = blockIdx.x * RB + threadIdx.x; k < n; k += stride)
   {
      size_t i0 = (size_t)(k >> 6) * (12 * SB) + (k & (SB - 1));
      float2 a[12], g[12];
#pragma unroll
      for (int u = 0; u < 12; u++)
      {
         a[u] = s[i0 + u * SB];
         g[u] = c ? c[i0 + u * SB] : make_float2(0.0f, 0.0f);
      }
#pragma unroll
      for (int j = 0; j < NQD; j++)
      {
         if (j < nq)
         {
            float2 b[12];
            const float2 *q = ql.q[j];
#pragma unroll
            for (int u = 0; u < 12; u++) b[u] = q[i0 + u * SB];
            pdot_reg(b, a, false, &re[j], &im[j]);
            if (c) pdot_reg(b, g, false, &gr[j], &gi[j]);
         }
      }
   }
   // work-group sums, one value at a time (few registers in flight)
   __shared__ dd sh[RB / 64][4 * NQD];
   const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
   for (int j = 0; j < NQD; j++)
   {
      const double src[4] = {re[j], im[j], gr[j], gi[j]};
#pragma unroll
      for (int u = 0; u < 4; u++)
      {
         dd x = {src[u], 0.0};
         for (int m = 32; m >= 1; m >>= 1) x = dd_add(x, shfl_xor_dd(x, m));
         if (l == 0) sh[w][4 * j + u] = x;
      }
   }
   __syncthreads();
   const int t = threadIdx.x;
   if (t < 4 * nq)
   {
      dd x = sh[0][t];
      for (int ww = 1; ww < RB / 64; ww++) x = dd_add(x, sh[ww][t]);
      part[(size_t)blockIdx.x * 4 * NQD + t] = x;
   }
}

// one work-group per vector j of the pass: p_j = coef[j0 + j]; the Gram
// entry G_{kn,j} = conj((q_j, q_kn)) of the newest vector kn goes to
// gram[kn * ld + j0 + j] if j0 + j < kn
__global__ __launch_bounds__(RB) void k_mdot_final(const dd *part, int nb, float2 *coef, int j0, float2 *gram, int ld,
                                                   int kn)
{
   const int j = blockIdx.x;
   dd v[4] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
   for (int k = threadIdx.x; k < nb; k += RB)
#pragma unroll
      for (int u = 0; u < 4; u++) v[u] = dd_add(v[u], part[(size_t)k * 4 * NQD + 4 * j + u]);
   block_reduce_dd<4>(v);
   if (threadIdx.x == 0)
   {
      coef[j0 + j] = make_float2((float)(v[0].hi + v[0].lo), (float)(v[1].hi + v[1].lo));
      if (kn > 0 && j0 + j < kn)
         gram[(size_t)kn * ld + j0 + j] = make_float2((float)(v[2].hi + v[2].lo), -(float)(v[3].hi + v[3].lo));
   }
}

// d_i = p_i - sum_{l<i} G_il d_l (in place in coef[0..k-1])
__global__ void k_gram_solve(float2 *coef, const float2 *gram, int ld, int k)
{
   if (threadIdx.x != 0) return;
   for (int i = 1; i < k; i++)
   {
      double zr = coef[i].x, zi = coef[i].y;
      for (int l = 0; l < i; l++)
      {
         float2 g = gram[(size_t)i * ld + l], d = coef[l];
         zr -= (double)g.x * d.x - (double)g.y * d.y;
         zi -= (double)g.x * d.y + (double)g.y * d.x;
      }
      coef[i] = make_float2((float)zr, (float)zi);
   }
}

// s -= sum_j coef[j0 + j] q_j (j < nq); in the last pass (fin) also the
// partial sums of |s|^2 and (s, r)
__global__ __launch_bounds__(RB) void k_maxpy_partial(float2 *s, QList ql, int nq, const float2 *coef, int j0,
                                                      const float2 *__restrict__ r, int fin, int n, dd *part)
{
   __shared__ float2 z[NQA];
   if (threadIdx.x < nq) z[threadIdx.x] = coef[j0 + threadIdx.x];
   __syncthreads();
   double nr = 0.0, pr = 0.0, pi = 0.0;
   int stride = gridDim.x * RB;
   for (int k = blockIdx.x * RB + threadIdx.x; k < n; k += stride)
   {
      size_t i0 = (size_t)(k >> 6) * (12 * SB) + (k & (SB - 1));
      float2 a[12];
#pragma unroll
      for (int c = 0; c < 12; c++) a[c] = s[i0 + c * SB];
      for (int j = 0; j < nq; j++)
      {
         const float2 *q = ql.q[j];
         const float2 zj = z[j];
#pragma unroll
         for (int c = 0; c < 12; c++)
         {
            float2 w = q[i0 + c * SB];
            a[c] = make_float2(a[c].x - (zj.x * w.x - zj.y * w.y), a[c].y - (zj.x * w.y + zj.y * w.x));
         }
      }
      if (nq)
      {
#pragma unroll
         for (int c = 0; c < 12; c++) s[i0 + c * SB] = a[c];
      }
      if (fin)
      {
         float2 b[12];
#pragma unroll
         for (int c = 0; c < 12; c++) b[c] = r[i0 + c * SB];
         double ni = 0.0;
         pdot_reg(a, a, true, &nr, &ni);
         pdot_reg(a, b, false, &pr, &pi);
      }
   }
   if (!fin) return;
   dd v[3] = {{nr, 0.0}, {pr, 0.0}, {pi, 0.0}};
   block_reduce_dd<3>(v);
   if (threadIdx.x == 0)
      for (int j = 0; j < 3; j++) part[(size_t)blockIdx.x * 3 + j] = v[j];
}

// coef[ib] = (|s|, |s|^2) as normalize() (salg.c), coef[ic] = (s, r) / |s|
__global__ __launch_bounds__(RB) void k_maxpy_final(const dd *part, int nb, float2 *coef, int ib, int ic)
{
   dd v[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
   for (int k = threadIdx.x; k < nb; k += RB)
#pragma unroll
      for (int j = 0; j < 3; j++) v[j] = dd_add(v[j], part[(size_t)k * 3 + j]);
   block_reduce_dd<3>(v);
   if (threadIdx.x == 0)
   {
      float r = (float)(v[0].hi + v[0].lo);
      float b = (float)sqrt((double)r);
      coef[ib] = make_float2(b, r);
      float rb = b > 0.0f ? 1.0f / b : 0.0f;
      coef[ic] = make_float2((float)(v[1].hi + v[1].lo) * rb, (float)(v[2].hi + v[2].lo) * rb);
   }
}

// s *= 1 / coef[ib].x, r -= coef[ic] s, partial sums of |r|^2
__global__ __launch_bounds__(RB) void k_scale_axpy_partial(float2 *s, float2 *r, const float2 *coef, int ib, int ic,
                                                           int n, dd *part)
{
   const float sc = 1.0f / coef[ib].x;
   const float zr = -coef[ic].x, zi = -coef[ic].y;
   double nr = 0.0, ni = 0.0;
   int stride = gridDim.x * RB;
   for (int k = blockIdx.x * RB + threadIdx.x; k < n; k += stride)
   {
      size_t i0 = (size_t)(k >> 6) * (12 * SB) + (k & (SB - 1));
      float2 a[12], b[12];
#pragma unroll
      for (int c = 0; c < 12; c++)
      {
         float2 u = s[i0 + c * SB];
         a[c] = make_float2(sc * u.x, sc * u.y);
         s[i0 + c * SB] = a[c];
         float2 w = r[i0 + c * SB];
         b[c] = make_float2(w.x + (zr * a[c].x - zi * a[c].y), w.y + (zr * a[c].y + zi * a[c].x));
         r[i0 + c * SB] = b[c];
      }
      pdot_reg(b, b, true, &nr, &ni);
   }
   dd v[2] = {{nr, 0.0}, {0.0, 0.0}};
   block_reduce_dd<2>(v);
   if (threadIdx.x == 0)
   {
      part[2 * blockIdx.x] = v[0];
      part[2 * blockIdx.x + 1] = v[1];
   }
}

void gs_classical(float2 *s, float2 *const *q, int k, float2 *rho, float2 *coef, int ib, int ic, int ir, int n,
                  float2 *d_gram, int gram_ld)
{
   Lib &L = lib();
   if (!d_part) d_part = (dd *)dev_alloc(sizeof(dd) * 2 * (RMAXB + 8));
   if (!d_mpart) d_mpart = (dd *)dev_alloc(sizeof(dd) * 4 * NQD * (RMAXB + 8));
   if (!d_part || !d_mpart || !d_gram || k > gram_ld) return;
   const int nb = red_blocks(n);
   // products with s and the Gram row of the newest vector q[k-1]
   for (int j0 = 0; j0 < k; j0 += NQD)
   {
      QList ql = {};
      int nq = std::min(NQD, k - j0);
      for (int j = 0; j < nq; j++) ql.q[j] = q[j0 + j];
      const float2 *c = (k >= 2 && j0 < k - 1) ? q[k - 1] : nullptr;
      k_mdot_partial<<<nb, RB, 0, L.stream>>>(s, c, ql, nq, n, d_mpart);
      k_mdot_final<<<nq, RB, 0, L.stream>>>(d_mpart, nb, coef, j0, d_gram, gram_ld, c ? k - 1 : 0);
   }
   if (k > 1) k_gram_solve<<<1, 64, 0, L.stream>>>(coef, d_gram, gram_ld, k);
   for (int j0 = 0;; j0 += NQA)
   {
      QList ql = {};
      int nq = std::max(0, std::min(NQA, k - j0));
      for (int j = 0; j < nq; j++) ql.q[j] = q[j0 + j];
      int fin = (j0 + NQA >= k);
      k_maxpy_partial<<<nb, RB, 0, L.stream>>>(s, ql, nq, coef, j0, rho, fin, n, d_mpart);
      if (fin) break;
   }
   k_maxpy_final<<<1, RB, 0, L.stream>>>(d_mpart, nb, coef, ib, ic);
   k_scale_axpy_partial<<<nb, RB, 0, L.stream>>>(s, rho, coef, ib, ic, n, d_part);
   k_gs_final<1><<<1, RB, 0, L.stream>>>(d_part, nb, coef, ir);
}

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------

static int nsites(int vol, int icom)
{
   Lib &L = lib();
   return (icom & 0x2) ? vol * L.g.nthread : vol;
}

static unsigned nb256(int n) { return (unsigned)((n + 255) / 256); }

struct Arg2
{
   Res a, b;
   int n;
   bool ok;
};

static Arg2 args2(const void *a, const void *b, int prec, int vol, int icom, const char *where,
                  bool upload_b = true)
{
   Arg2 g;
   g.ok = false;
   if (!require_geometry(where)) return g;
   g.n = nsites(vol, icom);
   if (g.n <= 0) return g;
   g.a = resolve(a, prec, g.n, 2, true);
   if (b) g.b = resolve(b, prec, g.n, 3, upload_b);
   g.ok = g.a.dev != nullptr && (!b || g.b.dev != nullptr);
   return g;
}

template <typename C, int K>
static void axpy(void *s, void *r, int prec, int vol, int icom, double zr, double zi, double cr, const char *where)
{
   Lib &L = lib();
   Arg2 g = args2(s, r, prec, vol, icom, where);
   if (!g.ok) return;
   typedef decltype(C::x) R;
   k_axpy<C, K><<<nb256(g.n), 256, 0, L.stream>>>((C *)g.a.dev, r ? (const C *)g.b.dev : nullptr, g.a.off,
                                                  r ? g.b.off : 0, g.n, (R)zr, (R)zi, (R)cr);
   finish(g.a, prec, g.n, true);
   if (r) finish(g.b, prec, g.n, false);
}

template <typename C, int KIND>
static void dot(void *s, void *r, int prec, int vol, int icom, dd *res, const char *where)
{
   Arg2 g = args2(s, r, prec, vol, icom, where);
   res[0] = res[1] = dd{0.0, 0.0};
   if (!g.ok) return;
   dot_sync<C, KIND>((const C *)g.a.dev, r ? (const C *)g.b.dev : (const C *)g.a.dev, g.a.off,
                     r ? g.b.off : g.a.off, g.n, res);
}

}  // namespace oqh

using namespace oqh;

// ---------------------------------------------------------------------------
// C ABI: linalg/salg.c (single precision)
// ---------------------------------------------------------------------------

extern "C" complex spinor_prod(int vol, int icom, spinor *s, spinor *r)
{
   dd res[2];
   dot<float2, DOT_C>(s, r, F32, vol, icom, res, "spinor_prod [salg.c]");
   complex z;
   z.re = (float)(res[0].hi + res[0].lo);
   z.im = (float)(res[1].hi + res[1].lo);
   return z;
}

extern "C" float spinor_prod_re(int vol, int icom, spinor *s, spinor *r)
{
   dd res[2];
   dot<float2, DOT_RE>(s, r, F32, vol, icom, res, "spinor_prod_re [salg.c]");
   return (float)(res[0].hi + res[0].lo);
}

extern "C" float norm_square(int vol, int icom, spinor *s)
{
   dd res[2];
   dot<float2, DOT_NORM>(s, nullptr, F32, vol, icom, res, "norm_square [salg.c]");
   return (float)(res[0].hi + res[0].lo);
}

extern "C" void mulc_spinor_add(int vol, int icom, spinor *s, spinor *r, complex z)
{
   axpy<float2, AX_MULC>(s, r, F32, vol, icom, z.re, z.im, 0.0, "mulc_spinor_add [salg.c]");
}

extern "C" void mulr_spinor_add(int vol, int icom, spinor *s, spinor *r, float c)
{
   axpy<float2, AX_MULR>(s, r, F32, vol, icom, c, 0.0, 0.0, "mulr_spinor_add [salg.c]");
}

extern "C" void scale(int vol, int icom, float c, spinor *s)
{
   axpy<float2, AX_SCALE>(s, nullptr, F32, vol, icom, c, 0.0, 0.0, "scale [salg.c]");
}

extern "C" void project(int vol, int icom, spinor *s, spinor *r)
{
   complex z = spinor_prod(vol, icom, r, s);
   z.re = -z.re;
   z.im = -z.im;
   mulc_spinor_add(vol, icom, s, r, z);
}

extern "C" float normalize(int vol, int icom, spinor *s)
{
   float r = norm_square(vol, icom, s);
   r = (float)sqrt((double)r);
   if (r != 0.0f)
      scale(vol, icom, 1.0f / r, s);
   else if (icom != 0)
      fail("normalize [salg.c]", "Vector has vanishing norm");
   return r;
}

extern "C" void mulg5(int vol, int icom, spinor *s)
{
   axpy<float2, AX_G5>(s, nullptr, F32, vol, icom, 0.0, 0.0, 0.0, "mulg5 [salg.c]");
}

extern "C" void mulmg5(int vol, int icom, spinor *s)
{
   axpy<float2, AX_MG5>(s, nullptr, F32, vol, icom, 0.0, 0.0, 0.0, "mulmg5 [salg.c]");
}

// ---------------------------------------------------------------------------
// C ABI: linalg/salg_dble.c (double precision, double-double sums)
// ---------------------------------------------------------------------------

static qflt to_qflt(dd v)
{
   qflt q;
   q.q[0] = v.hi;
   q.q[1] = v.lo;
   return q;
}

extern "C" complex_qflt spinor_prod_dble(int vol, int icom, spinor_dble *s, spinor_dble *r)
{
   dd res[2];
   dot<double2, DOT_C>(s, r, F64, vol, icom, res, "spinor_prod_dble [salg_dble.c]");
   complex_qflt z;
   z.re = to_qflt(res[0]);
   z.im = to_qflt(res[1]);
   return z;
}

extern "C" qflt spinor_prod_re_dble(int vol, int icom, spinor_dble *s, spinor_dble *r)
{
   dd res[2];
   dot<double2, DOT_RE>(s, r, F64, vol, icom, res, "spinor_prod_re_dble [salg_dble.c]");
   return to_qflt(res[0]);
}

extern "C" complex_qflt spinor_prod5_dble(int vol, int icom, spinor_dble *s, spinor_dble *r)
{
   dd res[2];
   dot<double2, DOT_C5>(s, r, F64, vol, icom, res, "spinor_prod5_dble [salg_dble.c]");
   complex_qflt z;
   z.re = to_qflt(res[0]);
   z.im = to_qflt(res[1]);
   return z;
}

extern "C" qflt norm_square_dble(int vol, int icom, spinor_dble *s)
{
   dd res[2];
   dot<double2, DOT_NORM>(s, nullptr, F64, vol, icom, res, "norm_square_dble [salg_dble.c]");
   return to_qflt(res[0]);
}

extern "C" void mulc_spinor_add_dble(int vol, int icom, spinor_dble *s, spinor_dble *r, complex_dble z)
{
   axpy<double2, AX_MULC>(s, r, F64, vol, icom, z.re, z.im, 0.0, "mulc_spinor_add_dble [salg_dble.c]");
}

extern "C" void mulr_spinor_add_dble(int vol, int icom, spinor_dble *s, spinor_dble *r, double c)
{
   axpy<double2, AX_MULR>(s, r, F64, vol, icom, c, 0.0, 0.0, "mulr_spinor_add_dble [salg_dble.c]");
}

extern "C" void combine_spinor_dble(int vol, int icom, spinor_dble *s, spinor_dble *r, double cs, double cr)
{
   axpy<double2, AX_COMB>(s, r, F64, vol, icom, cs, 0.0, cr, "combine_spinor_dble [salg_dble.c]");
}

extern "C" void scale_dble(int vol, int icom, double c, spinor_dble *s)
{
   axpy<double2, AX_SCALE>(s, nullptr, F64, vol, icom, c, 0.0, 0.0, "scale_dble [salg_dble.c]");
}

extern "C" void project_dble(int vol, int icom, spinor_dble *s, spinor_dble *r)
{
   complex_qflt q = spinor_prod_dble(vol, icom, r, s);
   complex_dble z;
   z.re = -q.re.q[0];
   z.im = -q.im.q[0];
   mulc_spinor_add_dble(vol, icom, s, r, z);
}

extern "C" double normalize_dble(int vol, int icom, spinor_dble *s)
{
   qflt q = norm_square_dble(vol, icom, s);
   double r = sqrt(q.q[0]);
   if (r != 0.0)
      scale_dble(vol, icom, 1.0 / r, s);
   else
      fail("normalize_dble [salg_dble.c]", "Vector field has vanishing norm");
   return r;
}

extern "C" void mulg5_dble(int vol, int icom, spinor_dble *s)
{
   axpy<double2, AX_G5>(s, nullptr, F64, vol, icom, 0.0, 0.0, 0.0, "mulg5_dble [salg_dble.c]");
}

extern "C" void mulmg5_dble(int vol, int icom, spinor_dble *s)
{
   axpy<double2, AX_MG5>(s, nullptr, F64, vol, icom, 0.0, 0.0, 0.0, "mulmg5_dble [salg_dble.c]");
}

// ---------------------------------------------------------------------------
// C ABI: utils/qsum.c -- host arithmetic on (hi,lo) pairs
// ---------------------------------------------------------------------------

extern "C" void acc_qflt(double u, double *qr)
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

extern "C" void add_qflt(double *qu, double *qv, double *qr)
{
   dd r = dd_add(dd{qu[0], qu[1]}, dd{qv[0], qv[1]});
   qr[0] = r.hi;
   qr[1] = r.lo;
}

// Dekker's exact product with the splitting constant 2^27+1 (qsum.c:68,313)
static void exact_prod(double u, double v, double *qr)
{
   const double mm = 134217729.0;
   double ux = mm * u, vx = mm * v;
   double uy = ux - u, vy = vx - v;
   double uh = ux - uy, vh = vx - vy;
   double ul = u - uh, vl = v - vh;
   double x = u * v;
   double c = uh * vh - x;
   double d = uh * vl + c;
   double e = ul * vh + d;
   double y = ul * vl + e;
   qr[0] = x;
   qr[1] = y;
}

extern "C" void scl_qflt(double u, double *qr)
{
   double p[2], s;
   exact_prod(u, qr[0], p);
   s = u * qr[1];
   acc_qflt(s, p);
   qr[0] = p[0];
   qr[1] = p[1];
}

extern "C" void mul_qflt(double *qu, double *qv, double *qr)
{
   double p[2], s;
   s = qu[0] * qv[1] + qu[1] * qv[0];
   exact_prod(qu[0], qv[0], p);
   acc_qflt(s, p);
   qr[0] = p[0];
   qr[1] = p[1];
}

extern "C" void global_qsum(int n, double **qu, double **qr)
{
   // one process: the global sum is the local value (qsum.c:416-423)
   for (int k = 0; k < n; k++)
   {
      qr[k][0] = qu[k][0];
      qr[k][1] = qu[k][1];
   }
}

// ---------------------------------------------------------------------------
// C ABI: sflds/sflds.c, sflds/unorm.c
// ---------------------------------------------------------------------------

extern "C" void set_s2zero(int vol, int icom, spinor *s)
{
   axpy<float2, AX_ZERO>(s, nullptr, F32, vol, icom, 0, 0, 0, "set_s2zero [sflds.c]");
}

extern "C" void set_sd2zero(int vol, int icom, spinor_dble *sd)
{
   axpy<double2, AX_ZERO>(sd, nullptr, F64, vol, icom, 0, 0, 0, "set_sd2zero [sflds.c]");
}

extern "C" void assign_s2s(int vol, int icom, spinor *s, spinor *r)
{
   axpy<float2, AX_COPY>(r, s, F32, vol, icom, 0, 0, 0, "assign_s2s [sflds.c]");
}

extern "C" void assign_sd2sd(int vol, int icom, spinor_dble *sd, spinor_dble *rd)
{
   axpy<double2, AX_COPY>(rd, sd, F64, vol, icom, 0, 0, 0, "assign_sd2sd [sflds.c]");
}

extern "C" void add_s2s(int vol, int icom, spinor *s, spinor *r)
{
   axpy<float2, AX_ADD>(r, s, F32, vol, icom, 0, 0, 0, "add_s2s [sflds.c]");
}

extern "C" void diff_s2s(int vol, int icom, spinor *s, spinor *r)
{
   // r = s - r
   axpy<float2, AX_DIFF>(r, s, F32, vol, icom, 0, 0, 0, "diff_s2s [sflds.c]");
}

template <int K>
static void mixed(const void *a, int pa, const void *b, void *r, int pr, int vol, int icom, const char *where)
{
   Lib &L = lib();
   if (!require_geometry(where)) return;
   int n = nsites(vol, icom);
   if (n <= 0) return;
   Res ra = resolve(a, pa, n, 0, true);
   Res rb;
   if (b) rb = resolve(b, pa, n, 1, true);
   Res rr = resolve(r, pr, n, 2, K == 2);
   if (!ra.dev || !rr.dev || (b && !rb.dev)) return;
   k_mixed<K><<<nb256(n), 256, 0, L.stream>>>(ra.dev, b ? rb.dev : nullptr, rr.dev, ra.off, b ? rb.off : 0,
                                              rr.off, n);
   finish(rr, pr, n, true);
}

extern "C" void assign_s2sd(int vol, int icom, spinor *s, spinor_dble *rd)
{
   mixed<0>(s, F32, nullptr, rd, F64, vol, icom, "assign_s2sd [sflds.c]");
}

extern "C" void assign_sd2s(int vol, int icom, spinor_dble *sd, spinor *r)
{
   mixed<1>(sd, F64, nullptr, r, F32, vol, icom, "assign_sd2s [sflds.c]");
}

extern "C" void add_s2sd(int vol, int icom, spinor *s, spinor_dble *rd)
{
   mixed<2>(s, F32, nullptr, rd, F64, vol, icom, "add_s2sd [sflds.c]");
}

extern "C" void diff_sd2s(int vol, int icom, spinor_dble *sd, spinor_dble *rd, spinor *r)
{
   mixed<3>(sd, F64, rd, r, F32, vol, icom, "diff_sd2s [sflds.c]");
}

template <typename C> static double unorm_any(const void *s, int prec, int vol, int icom, const char *where)
{
   Lib &L = lib();
   if (!require_geometry(where)) return 0.0;
   int n = nsites(vol, icom);
   if (n <= 0) return 0.0;
   Res r = resolve(s, prec, n, 0, true);
   if (!r.dev) return 0.0;
   static double *d_p = nullptr, *h_p = nullptr;
   if (!d_p) d_p = (double *)dev_alloc(sizeof(double) * (RMAXB + 8));
   if (!h_p) HIP_OK(hipHostMalloc((void **)&h_p, sizeof(double), 0));
   int nb = red_blocks(n);
   k_unorm_partial<C><<<nb, RB, 0, L.stream>>>((const C *)r.dev, r.off, n, d_p);
   k_max_final<<<1, 64, 0, L.stream>>>(d_p, nb, d_p + RMAXB);
   HIP_OK(hipMemcpyAsync(h_p, d_p + RMAXB, sizeof(double), hipMemcpyDeviceToHost, L.stream));
   HIP_OK(hipStreamSynchronize(L.stream));
   return *h_p;
}

extern "C" float unorm(int vol, int icom, spinor *s)
{
   double m = unorm_any<float2>(s, F32, vol, icom, "unorm [unorm.c]");
   return (float)sqrt((double)(float)m);
}

extern "C" double unorm_dble(int vol, int icom, spinor_dble *sd)
{
   return sqrt(unorm_any<double2>(sd, F64, vol, icom, "unorm_dble [unorm.c]"));
}

// ---------------------------------------------------------------------------
// spinor_prod_dble without allocation (no counterpart in the reference)
//
// For callers that cannot allocate device memory (code under HIP graph
// capture, tight loops, harnesses that forbid hipMalloc).  The reduction's
// work-group partials live in the file-scope array scratch (RMAXB work-groups
// x two double-double numbers = 32 KiB per GPU, held until the process
// ends).  The host name of the array is not a device address: its address on
// a device is looked up once per device with hipGetSymbolAddress and only
// that pointer is used.  Every partial the final kernel reads is written by
// the same call before, so nothing needs clearing; calls that overlap on one
// GPU share the array and race.
// ---------------------------------------------------------------------------

namespace oqh {

constexpr int NOALLOC_MAXDEV = 64;
static __device__ __align__(256) double scratch[4 * RMAXB];
static dd *scratch_ptr[NOALLOC_MAXDEV] = {};

// scratch of the library's device, looked up on the first call there
static dd *noalloc_scratch(const char *where)
{
   Lib &L = lib();
   int cur = -1;
   if (hipGetDevice(&cur) != hipSuccess || cur != L.device)
   {
      fail(where, "the current HIP device must be the library's (oqh_init)");
      return nullptr;
   }
   if (cur < 0 || cur >= NOALLOC_MAXDEV)
   {
      fail(where, "device number out of range");
      return nullptr;
   }
   if (!scratch_ptr[cur])
   {
      void *p = nullptr;
      hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || p == nullptr)
      {
         fail(where, e != hipSuccess ? hipGetErrorString(e) : "no device address of the scratch array");
         return nullptr;
      }
      scratch_ptr[cur] = (dd *)p;
   }
   return scratch_ptr[cur];
}

}  // namespace oqh

extern "C" int oqh_spinor_prod_dble_noalloc(int vol, int icom, spinor_dble *s, spinor_dble *r, complex_qflt *z)
{
   const char *where = "oqh_spinor_prod_dble_noalloc";
   Lib &L = lib();
   if (!L.inited || L.g.V == 0)
   {
      fail(where, "the library has no device or no lattice geometry (oqh_init, oqh_geometry)");
      return 1;
   }
   dd *part = noalloc_scratch(where);
   if (!part) return 1;
   const int n = nsites(vol, icom);
   if (n <= 0)
   {
      fail(where, "vol must be positive");
      return 1;
   }
   int os = 0, orr = 0, ns = 0, nr = 0;
   const double2 *ds = (const double2 *)device_field_n(s, F64, &os, &ns);
   const double2 *dr = (const double2 *)device_field_n(r, F64, &orr, &nr);
   if (!ds || !dr)
   {
      fail(where, "s and r must be double-precision device fields of the library (alloc_wsd, oqh_alloc_sd)");
      return 1;
   }
   if (n > ns || n > nr)
   {
      fail(where, "vol exceeds the points of the fields");
      return 1;
   }
   hipPointerAttribute_t at;
   if (z == nullptr || hipPointerGetAttributes(&at, z) != hipSuccess || at.type != hipMemoryTypeDevice)
   {
      (void)hipGetLastError();
      fail(where, "z must point to device memory");
      return 1;
   }
   const int nb = red_blocks(n);  // at most RMAXB partials: every vol fits the scratch
   if (nb > RMAXB)
   {
      fail(where, "too many work-group partials for the scratch array");
      return 1;
   }
   k_dot_partial<double2, DOT_C><<<nb, RB, 0, L.stream>>>(ds, dr, os, orr, n, part);
   k_dd_final<<<1, RB, 0, L.stream>>>(part, nb, (dd *)z);
   return 0;
}
