// ptc.hip -- gauge-equivariant network used as a learned preconditioner of
// sap_gcr (no counterpart in the reference).
//
// A field with C channels holds C colour vectors per point in the unit
// layout of the spinor fields (unit 3 i + a of point x is channel i, colour
// a; a spinor is the field with its 4 spin components as channels).  One
// layer maps C_in to C_out channels through the parallel-transport stencil
//
//    out(x)[o] = sum_{p=0..8} sum_i W[p][o][i] (T_p in)(x)[i],
//    T_0 = 1,  T_{2mu+1} in(x) = U(x,mu) in(x+mu),
//    T_{2mu+2} in(x) = U(x-mu,mu)^+ in(x-mu),
//
// with complex weights acting on the channel index only (the layer commutes
// with gauge transformations).  The network is linear: a learned
// approximation of D^-1 applied to the residual that the Schwarz cycles
// leave.
//
// k_ptc: a work-group takes 128 points.  For two directions p at a time its
// threads transport the input channels of the neighbours with the links (FP32
// in training, an FP16 copy in the preconditioner, where the links are most
// of the bytes a layer reads) and store them as BF16 in LDS (a K x N matrix, K = real and imaginary
// parts of the C_in channels, N = 3 colours x 128 points); the channel
// mixing is then a real GEMM [2 C_out x K] x [K x N] on the matrix cores
// (v_mfma_f32_32x32x16_bf16, FP32 accumulation over the 9 directions), the
// weights in the real form [[Re W, -Im W], [Im W, Re W]] as BF16.  The
// input gradient of a layer is the same kernel with the weights
// W'[p] = W[pbar]^+ (T_p^+ = T_pbar); the weight gradient
// dW[p][o][i] = sum_{x,a} g(x)[o][a] conj((T_p in)(x)[i][a]) is formed in
// FP32 per work-group and summed in double.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "dirac_dev.h"
#include "oqh_internal.h"

namespace oqh {

constexpr int NDIR = 9;
constexpr int PNS = 128;  // points per work-group
constexpr int PWG = 256;  // threads per work-group
constexpr int NLMAX = 8;  // layers
constexpr int CMAX = 16;  // channels

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// K of one direction (2 C reals) rounded up to the MFMA depth 16
__host__ __device__ constexpr int kpad(int c) { return ((2 * c + 15) / 16) * 16; }

__device__ inline int pbar_of(int p) { return p == 0 ? 0 : ((p & 1) ? p + 1 : p - 1); }

// address of the link of the transport T_p at point x (neighbour y), and
// whether it enters daggered
template <typename UL>
__device__ inline const UL *ptc_link(const UL *u, int x, int y, int p, int h, bool &dag)
{
   const int mu = (p - 1) >> 1;
   const bool fwd = p & 1;
   dag = !fwd;
   if (fwd) return (x >= h) ? link_ptr(u, x - h, 2 * mu) : link_ptr(u, y - h, 2 * mu + 1);
   return (x >= h) ? link_ptr(u, x - h, 2 * mu + 1) : link_ptr(u, y - h, 2 * mu);
}

__device__ inline int ptc_nbr(const int *nbr, int V, int x, int p)
{
   if (p == 0) return x;
   const int mu = (p - 1) >> 1;
   return nbr[(size_t)((p & 1) ? 2 * mu : 2 * mu + 1) * V + x] & 0x7fffffff;
}

// (T_p in)(x) for channel i: w[3]
template <int CI>
__device__ inline void ptc_transport(const float2 *in, int y, int i, int p, const float2 *um, bool dag, float2 *w)
{
   float2 v[3];
#pragma unroll
   for (int a = 0; a < 3; a++) v[a] = in[unit_index(y, 3 * i + a, 3 * CI)];
   if (p == 0)
   {
#pragma unroll
      for (int a = 0; a < 3; a++) w[a] = v[a];
   }
   else if (dag)
      su3_mul<float2, true>(um, v, w);
   else
      su3_mul<float2, false>(um, v, w);
}

// activations: FP32 complex numbers (training, the network's input and
// output), or BF16 pairs between the layers of the preconditioner
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

__device__ inline float2 ld_act(const float2 *p, size_t i) { return p[i]; }
__device__ inline float2 ld_act(const bf16x2 *p, size_t i)
{
   const bf16x2 b = p[i];
   return make_float2((float)b.x, (float)b.y);
}
__device__ inline void st_act(float2 *p, size_t i, float2 v) { p[i] = v; }
__device__ inline void st_act(bf16x2 *p, size_t i, float2 v)
{
   bf16x2 b;
   b.x = (__bf16)v.x;
   b.y = (__bf16)v.y;
   p[i] = b;
}

// links: FP32, or the FP16 copy that the preconditioner reads
__device__ inline float2 lnk(float2 v) { return v; }
__device__ inline float2 lnk(__half2 v) { return __half22float2(v); }

// out (+)= layer(in); wp: packed BF16 weights [9][32][kpad(CI)].  Two
// directions per phase (five phases): the two halves of the work-group
// transport all input channels of their point for direction 2q and 2q+1,
// so every link is loaded once, and one barrier pair serves two directions.
// accum (out += ...) only with FP32 output.
template <int CI, int CO, typename TI = float2, typename TO = float2, typename UL = float2>
__global__ __launch_bounds__(PWG) void k_ptc(const TI *__restrict__ in, TO *__restrict__ out,
                                             const __bf16 *__restrict__ wp, const UL *__restrict__ u,
                                             const int *__restrict__ nbr, int V, int accum)
{
   constexpr int KP = kpad(CI);
   constexpr int LD = 2 * KP + 8;  // LDS row stride (bf16): two direction blocks
   static_assert(2 * CO <= 32 && CI % 2 == 0, "channel counts");
   __shared__ __attribute__((aligned(16))) __bf16 Bt[3 * PNS * LD];
   const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
   const int x0 = blockIdx.x * PNS, h = V / 2;
   for (int q = t; q < 3 * PNS; q += PWG)
      for (int k = 2 * CI; k < KP; k++) Bt[q * LD + k] = Bt[q * LD + KP + k] = (__bf16)0.0f;
   f32x16 acc[3];
#pragma unroll
   for (int tt = 0; tt < 3; tt++)
#pragma unroll
      for (int j = 0; j < 16; j++) acc[tt][j] = 0.0f;
   const int s = t & (PNS - 1), hh = t >> 7;
   const int x = x0 + s;
   for (int ph = 0; ph < (NDIR + 1) / 2; ph++)
   {
      const int p = 2 * ph + hh;
      __bf16 *brow = Bt + s * LD + hh * KP;
      if (x < V && p < NDIR)
      {
         const int y = ptc_nbr(nbr, V, x, p);
         float2 um[9];
         bool dag = false;
         if (p > 0)
         {
            const UL *up = ptc_link(u, x, y, p, h, dag);
#pragma unroll
            for (int k = 0; k < 9; k++) um[k] = lnk(up[k * SB]);
         }
         // all loads first (independent), then the products
         float2 v[CI][3];
#pragma unroll
         for (int i = 0; i < CI; i++)
#pragma unroll
            for (int a = 0; a < 3; a++) v[i][a] = ld_act(in, unit_index(y, 3 * i + a, 3 * CI));
#pragma unroll
         for (int i = 0; i < CI; i++)
         {
            float2 w[3];
            if (p == 0)
            {
#pragma unroll
               for (int a = 0; a < 3; a++) w[a] = v[i][a];
            }
            else if (dag)
               su3_mul<float2, true>(um, v[i], w);
            else
               su3_mul<float2, false>(um, v[i], w);
#pragma unroll
            for (int a = 0; a < 3; a++)
            {
               bf16x2 b;
               b.x = (__bf16)w[a].x;
               b.y = (__bf16)w[a].y;
               *(bf16x2 *)(brow + a * PNS * LD + 2 * i) = b;
            }
         }
      }
      else
      {
         bf16x2 z;
         z.x = z.y = (__bf16)0.0f;
#pragma unroll
         for (int i = 0; i < CI; i++)
#pragma unroll
            for (int a = 0; a < 3; a++) *(bf16x2 *)(brow + a * PNS * LD + 2 * i) = z;
      }
      __syncthreads();
#pragma unroll
      for (int tt = 0; tt < 3; tt++)
      {
         const int n = (wv + 4 * tt) * 32 + (lane & 31);
#pragma unroll
         for (int d = 0; d < 2; d++)
         {
            const int pd = 2 * ph + d;
            if (pd >= NDIR) break;
#pragma unroll
            for (int k0 = 0; k0 < KP; k0 += 16)
            {
               const bf16x8 a = *(const bf16x8 *)(wp + (size_t)(pd * 32 + (lane & 31)) * KP + k0 + 8 * (lane >> 5));
               const bf16x8 b = *(const bf16x8 *)(Bt + n * LD + d * KP + k0 + 8 * (lane >> 5));
               acc[tt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[tt], 0, 0, 0);
            }
         }
      }
      __syncthreads();
   }
   // D[m][n]: lane holds n = lane % 32 and rows 8 (j/4) + 4 (lane/32) + j%4;
   // rows 2o, 2o+1 are Re and Im of channel o
#pragma unroll
   for (int tt = 0; tt < 3; tt++)
   {
      const int n = (wv + 4 * tt) * 32 + (lane & 31);
      const int a = n / PNS, xo = x0 + n % PNS;
      if (xo >= V) continue;
#pragma unroll
      for (int j = 0; j < 16; j += 2)
      {
         const int m = 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3);
         const int o = m >> 1;
         if (o < CO)
         {
            float2 v = make_float2(acc[tt][j], acc[tt][j + 1]);
            const size_t id = unit_index(xo, 3 * o + a, 3 * CO);
            if (accum)
            {
               float2 w = ld_act(out, id);
               v.x += w.x;
               v.y += w.y;
            }
            st_act(out, id, v);
         }
      }
   }
}

// packed BF16 weights of a layer with ci inputs and co outputs from the
// complex weights w[9][co][ci]; adj: the adjoint layer (co inputs, ci
// outputs) with W'[p] = W[pbar]^+
__global__ void k_ptc_pack(const float2 *w, int ci, int co, int adj, __bf16 *wp)
{
   const int p = blockIdx.x;
   const int cin = adj ? co : ci, cout = adj ? ci : co;
   const int KP = kpad(cin);
   for (int e = threadIdx.x; e < 32 * KP; e += blockDim.x)
   {
      const int m = e / KP, k = e % KP;
      const int o = m >> 1, r = m & 1, i = k >> 1, q = k & 1;
      float val = 0.0f;
      if (o < cout && i < cin)
      {
         float2 z;
         if (!adj)
            z = w[((size_t)p * co + o) * ci + i];
         else
         {
            float2 y = w[((size_t)pbar_of(p) * co + i) * ci + o];
            z = make_float2(y.x, -y.y);
         }
         // [[Re, -Im], [Im, Re]]
         val = (r == q) ? z.x : (r == 0 ? -z.y : z.y);
      }
      wp[((size_t)p * 32 + m) * KP + k] = (__bf16)val;
   }
}

// partial weight gradients of one work-group: part[blk][p][o][i]
template <int CI, int CO>
__global__ __launch_bounds__(PWG) void k_ptc_wgrad(const float2 *__restrict__ in, const float2 *__restrict__ g,
                                                   const float2 *__restrict__ u, const int *__restrict__ nbr, int V,
                                                   float2 *part)
{
   __shared__ float2 T[3 * CI][PNS];
   __shared__ float2 Gs[3 * CO][PNS];
   const int t = threadIdx.x, x0 = blockIdx.x * PNS, h = V / 2;
   for (int e = t; e < 3 * CO * PNS; e += PWG)
   {
      const int c = e / PNS, s = e % PNS;
      Gs[c][s] = (x0 + s < V) ? g[unit_index(x0 + s, c, 3 * CO)] : make_float2(0.0f, 0.0f);
   }
   const int s = t & (PNS - 1), hh = t >> 7;
   const int x = x0 + s;
   constexpr int CH = CI / 2;
   for (int p = 0; p < NDIR; p++)
   {
      if (x < V)
      {
         const int y = ptc_nbr(nbr, V, x, p);
         float2 um[9];
         bool dag = false;
         if (p > 0)
         {
            const float2 *up = ptc_link(u, x, y, p, h, dag);
#pragma unroll
            for (int k = 0; k < 9; k++) um[k] = up[k * SB];
         }
#pragma unroll
         for (int c = 0; c < CH; c++)
         {
            const int i = hh * CH + c;
            float2 w[3];
            ptc_transport<CI>(in, y, i, p, um, dag, w);
#pragma unroll
            for (int a = 0; a < 3; a++) T[3 * i + a][s] = w[a];
         }
      }
      else
      {
         for (int c = 0; c < CH; c++)
            for (int a = 0; a < 3; a++) T[3 * (hh * CH + c) + a][s] = make_float2(0.0f, 0.0f);
      }
      __syncthreads();
      for (int e = t; e < CO * CI; e += PWG)
      {
         const int o = e / CI, i = e % CI;
         float zr = 0.0f, zi = 0.0f;
         for (int a = 0; a < 3; a++)
            for (int q = 0; q < PNS; q++)
            {
               const float2 gv = Gs[3 * o + a][q], tv = T[3 * i + a][q];
               // g conj(t)
               zr = fmaf(gv.x, tv.x, fmaf(gv.y, tv.y, zr));
               zi = fmaf(gv.y, tv.x, fmaf(-gv.x, tv.y, zi));
            }
         part[((size_t)blockIdx.x * NDIR + p) * CO * CI + e] = make_float2(zr, zi);
      }
      __syncthreads();
   }
}

// grad[e] += sum over work-groups of part (double)
__global__ void k_ptc_wgrad_sum(const float2 *part, int nb, int n, float2 *grad)
{
   const int e = blockIdx.x * blockDim.x + threadIdx.x;
   if (e >= n) return;
   double zr = 0.0, zi = 0.0;
   for (int b = 0; b < nb; b++)
   {
      float2 v = part[(size_t)b * n + e];
      zr += v.x;
      zi += v.y;
   }
   grad[e].x += (float)zr;
   grad[e].y += (float)zi;
}

// loss terms of one sample: partial sums of |e|^2 and |out - e|^2
__global__ void k_ptc_loss(const float2 *out, const float2 *e, int n, double *part)
{
   __shared__ double sh[2][256];
   double se = 0.0, sd = 0.0;
   for (int k = blockIdx.x * 256 + threadIdx.x; k < n; k += gridDim.x * 256)
   {
      float2 a = out[k], b = e[k];
      double dx = (double)a.x - b.x, dy = (double)a.y - b.y;
      se += (double)b.x * b.x + (double)b.y * b.y;
      sd += dx * dx + dy * dy;
   }
   sh[0][threadIdx.x] = se;
   sh[1][threadIdx.x] = sd;
   __syncthreads();
   for (int m = 128; m > 0; m >>= 1)
   {
      if (threadIdx.x < m)
      {
         sh[0][threadIdx.x] += sh[0][threadIdx.x + m];
         sh[1][threadIdx.x] += sh[1][threadIdx.x + m];
      }
      __syncthreads();
   }
   if (threadIdx.x == 0)
   {
      part[2 * blockIdx.x] = sh[0][0];
      part[2 * blockIdx.x + 1] = sh[1][0];
   }
}

// g = c (out - e)
__global__ void k_ptc_lgrad(const float2 *out, const float2 *e, float c, int n, float2 *g)
{
   const int k = blockIdx.x * 256 + threadIdx.x;
   if (k < n) g[k] = make_float2(c * (out[k].x - e[k].x), c * (out[k].y - e[k].y));
}

// Adam step on n reals
__global__ void k_adam(float *w, const float *g, float *m, float *v, int n, float lr, float b1, float b2, float eps,
                       float c1, float c2)
{
   const int k = blockIdx.x * 256 + threadIdx.x;
   if (k >= n) return;
   const float gk = g[k];
   m[k] = b1 * m[k] + (1.0f - b1) * gk;
   v[k] = b2 * v[k] + (1.0f - b2) * gk * gk;
   w[k] -= lr * (m[k] * c1) / (sqrtf(v[k] * c2) + eps);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

__global__ void k_ptc_link16(const float2 *a, __half2 *b, size_t n)
{
   const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
   if (i < n) b[i] = __float22half2_rn(a[i]);
}

namespace {

struct Net
{
   int nl = 0;
   int w[NLMAX + 1] = {};
   size_t off[NLMAX + 1] = {};           // offsets of the layers in the weight array
   float2 *d_w = nullptr, *d_g = nullptr;  // complex weights and gradients
   float *d_m = nullptr, *d_v = nullptr;   // Adam moments
   __bf16 *d_pf[NLMAX] = {}, *d_pa[NLMAX] = {};
   bool packed = false;
   int adam_t = 0;
   int V = 0;
   float2 *d_act[NLMAX + 1] = {};  // activations of the last forward
   float2 *d_gb[2] = {};            // backward gradients (ping-pong)
   float2 *d_part = nullptr;        // weight-gradient partials
   double *d_lpart = nullptr, *h_lpart = nullptr;
   float2 *d_tmp = nullptr;         // spinor work field of the preconditioner
   bf16x2 *d_inf[2] = {};           // BF16 activations of the preconditioner (ping-pong)
   __half2 *d_u16 = nullptr;        // FP16 links of the preconditioner
   int u16_epoch = -1;              // link epoch of d_u16
};

Net N;

bool valid_width(int c) { return c == 4 || c == 8 || c == 12 || c == 16; }

void free_net()
{
   dev_free(N.d_w); dev_free(N.d_g); dev_free(N.d_m); dev_free(N.d_v);
   for (int l = 0; l < NLMAX; l++) { dev_free(N.d_pf[l]); dev_free(N.d_pa[l]); }
   for (int l = 0; l <= NLMAX; l++) dev_free(N.d_act[l]);
   dev_free(N.d_gb[0]); dev_free(N.d_gb[1]); dev_free(N.d_part); dev_free(N.d_lpart); dev_free(N.d_tmp);
   dev_free(N.d_inf[0]); dev_free(N.d_inf[1]); dev_free(N.d_u16);
   if (N.h_lpart) HIP_OK(hipHostFree(N.h_lpart));
   N = Net();
}

size_t nweights() { return N.off[N.nl]; }

void pack()
{
   if (N.packed) return;
   Lib &L = lib();
   for (int l = 0; l < N.nl; l++)
   {
      k_ptc_pack<<<NDIR, 256, 0, L.stream>>>(N.d_w + N.off[l], N.w[l], N.w[l + 1], 0, N.d_pf[l]);
      k_ptc_pack<<<NDIR, 256, 0, L.stream>>>(N.d_w + N.off[l], N.w[l], N.w[l + 1], 1, N.d_pa[l]);
   }
   N.packed = true;
}

// activations sized for the current lattice
bool ensure_fields()
{
   Lib &L = lib();
   if (N.V == L.g.V && N.d_gb[0]) return true;
   for (int l = 0; l <= NLMAX; l++) { dev_free(N.d_act[l]); N.d_act[l] = nullptr; }
   for (int k = 0; k < 2; k++) { dev_free(N.d_gb[k]); N.d_gb[k] = nullptr; }
   for (int k = 0; k < 2; k++) { dev_free(N.d_inf[k]); N.d_inf[k] = nullptr; }
   dev_free(N.d_part); dev_free(N.d_tmp); dev_free(N.d_u16);
   N.d_u16 = nullptr;
   N.u16_epoch = -1;
   N.V = L.g.V;
   for (int l = 1; l < N.nl; l++) N.d_act[l] = (float2 *)dev_alloc(units_for(N.V, 3 * CMAX) * sizeof(float2));
   for (int k = 0; k < 2; k++) N.d_gb[k] = (float2 *)dev_alloc(units_for(N.V, 3 * CMAX) * sizeof(float2));
   const int nb = (N.V + PNS - 1) / PNS;
   N.d_part = (float2 *)dev_alloc((size_t)nb * NDIR * CMAX * CMAX * sizeof(float2));
   N.d_tmp = (float2 *)dev_alloc(units_for(N.V, 12) * sizeof(float2));
   for (int k = 0; k < 2; k++) N.d_inf[k] = (bf16x2 *)dev_alloc(units_for(N.V, 3 * CMAX) * sizeof(bf16x2));
   N.d_u16 = (__half2 *)dev_alloc(units_for(N.V / 2, 72) * sizeof(__half2));
   return N.d_gb[0] && N.d_gb[1] && N.d_part && N.d_tmp && N.d_inf[0] && N.d_inf[1] && N.d_u16;
}

template <int CI, int CO>
void launch_layer(const float2 *in, float2 *out, const __bf16 *wp, int accum)
{
   Lib &L = lib();
   const int V = L.g.V;
   k_ptc<CI, CO><<<(V + PNS - 1) / PNS, PWG, 0, L.stream>>>(in, out, wp, L.d_u, L.d_nbr, V, accum);
}

// one layer of the preconditioner: mode 0 FP32 -> BF16, 1 BF16 -> BF16,
// 2 BF16 -> FP32 added to out, 3 FP32 -> FP32 added to out
template <int CI, int CO>
void launch_inf(const void *in, void *out, const __bf16 *wp, int mode)
{
   Lib &L = lib();
   const int V = L.g.V;
   const unsigned nb = (V + PNS - 1) / PNS;
   const __half2 *u = N.d_u16;
   if (mode == 0)
      k_ptc<CI, CO, float2, bf16x2, __half2><<<nb, PWG, 0, L.stream>>>((const float2 *)in, (bf16x2 *)out, wp, u,
                                                                       L.d_nbr, V, 0);
   else if (mode == 1)
      k_ptc<CI, CO, bf16x2, bf16x2, __half2><<<nb, PWG, 0, L.stream>>>((const bf16x2 *)in, (bf16x2 *)out, wp, u,
                                                                       L.d_nbr, V, 0);
   else if (mode == 2)
      k_ptc<CI, CO, bf16x2, float2, __half2><<<nb, PWG, 0, L.stream>>>((const bf16x2 *)in, (float2 *)out, wp, u,
                                                                       L.d_nbr, V, 1);
   else
      k_ptc<CI, CO, float2, float2, __half2><<<nb, PWG, 0, L.stream>>>((const float2 *)in, (float2 *)out, wp, u,
                                                                       L.d_nbr, V, 1);
}

template <int CI, int CO> void launch_wgrad(const float2 *in, const float2 *g)
{
   Lib &L = lib();
   const int V = L.g.V, nb = (V + PNS - 1) / PNS;
   k_ptc_wgrad<CI, CO><<<nb, PWG, 0, L.stream>>>(in, g, L.d_u, L.d_nbr, V, N.d_part);
}

#define OQH_PTC_DISPATCH(F, ci, co, ...)                                             \
   do                                                                               \
   {                                                                                \
      const int ci_ = (ci), co_ = (co);                                             \
      bool done_ = false;                                                           \
      OQH_PTC_CASE(F, 4, 4, __VA_ARGS__) OQH_PTC_CASE(F, 4, 8, __VA_ARGS__)           \
      OQH_PTC_CASE(F, 4, 12, __VA_ARGS__) OQH_PTC_CASE(F, 4, 16, __VA_ARGS__)         \
      OQH_PTC_CASE(F, 8, 4, __VA_ARGS__) OQH_PTC_CASE(F, 8, 8, __VA_ARGS__)           \
      OQH_PTC_CASE(F, 8, 12, __VA_ARGS__) OQH_PTC_CASE(F, 8, 16, __VA_ARGS__)         \
      OQH_PTC_CASE(F, 12, 4, __VA_ARGS__) OQH_PTC_CASE(F, 12, 8, __VA_ARGS__)         \
      OQH_PTC_CASE(F, 12, 12, __VA_ARGS__) OQH_PTC_CASE(F, 12, 16, __VA_ARGS__)       \
      OQH_PTC_CASE(F, 16, 4, __VA_ARGS__) OQH_PTC_CASE(F, 16, 8, __VA_ARGS__)         \
      OQH_PTC_CASE(F, 16, 12, __VA_ARGS__) OQH_PTC_CASE(F, 16, 16, __VA_ARGS__)       \
      if (!done_) fail("ptc", "unsupported channel counts");                        \
   } while (0)
#define OQH_PTC_CASE(F, A, B, ...)                                                   \
   if (!done_ && ci_ == A && co_ == B)                                              \
   {                                                                                \
      F<A, B>(__VA_ARGS__);                                                         \
      done_ = true;                                                                 \
   }

// forward pass; keeps the activations of the hidden layers
void forward(const float2 *in, float2 *out)
{
   pack();
   const float2 *x = in;
   for (int l = 0; l < N.nl; l++)
   {
      float2 *y = (l == N.nl - 1) ? out : N.d_act[l + 1];
      OQH_PTC_DISPATCH(launch_layer, N.w[l], N.w[l + 1], x, y, N.d_pf[l], 0);
      x = y;
   }
}

// gradients of the weights (added to d_g) for the output gradient g of the
// last forward pass with input in
void backward(const float2 *in, const float2 *g)
{
   Lib &L = lib();
   const int nb = (L.g.V + PNS - 1) / PNS;
   const float2 *gl = g;
   for (int l = N.nl - 1; l >= 0; l--)
   {
      const float2 *al = l ? N.d_act[l] : in;
      OQH_PTC_DISPATCH(launch_wgrad, N.w[l], N.w[l + 1], al, gl);
      const int n = NDIR * N.w[l] * N.w[l + 1];
      k_ptc_wgrad_sum<<<(n + 255) / 256, 256, 0, L.stream>>>(N.d_part, nb, n, N.d_g + N.off[l]);
      if (l > 0)
      {
         float2 *gn = (gl == N.d_gb[0]) ? N.d_gb[1] : N.d_gb[0];
         OQH_PTC_DISPATCH(launch_layer, N.w[l + 1], N.w[l], gl, gn, N.d_pa[l], 0);
         gl = gn;
      }
   }
}

// mean loss over the samples; gradients of the mean loss in d_g
double loss_and_grad(int n, float2 *const *r, float2 *const *e, float2 *out)
{
   Lib &L = lib();
   const int V = L.g.V;
   const size_t nu = units_for(V, 12);
   HIP_OK(hipMemsetAsync(N.d_g, 0, nweights() * sizeof(float2), L.stream));
   double loss = 0.0;
   const int lb = 256;
   for (int k = 0; k < n; k++)
   {
      forward(r[k], out);
      k_ptc_loss<<<lb, 256, 0, L.stream>>>(out, e[k], (int)nu, N.d_lpart);
      HIP_OK(hipMemcpyAsync(N.h_lpart, N.d_lpart, 2 * lb * sizeof(double), hipMemcpyDeviceToHost, L.stream));
      HIP_OK(hipStreamSynchronize(L.stream));
      double se = 0.0, sd = 0.0;
      for (int b = 0; b < lb; b++)
      {
         se += N.h_lpart[2 * b];
         sd += N.h_lpart[2 * b + 1];
      }
      if (se == 0.0) continue;
      loss += sd / se / n;
      // g = 2 (out - e) / (|e|^2 n), into the ping-pong buffer not used by
      // the first backward step
      float2 *g = N.d_gb[1];
      k_ptc_lgrad<<<(unsigned)((nu + 255) / 256), 256, 0, L.stream>>>(out, e[k], (float)(2.0 / (se * n)), (int)nu, g);
      backward(r[k], g);
   }
   return loss;
}

bool require_net(const char *where)
{
   if (!N.nl)
   {
      fail(where, "the network is not created");
      return false;
   }
   if (!require_geometry(where) || !lib().ud_set)
   {
      if (lib().g.V) fail(where, "gauge field is not set");
      return false;
   }
   ensure_u();
   return ensure_fields();
}

// device fields of n spinors, all registered full fields
bool fields_of(int n, spinor **s, std::vector<float2 *> &out, const char *where)
{
   out.resize(n);
   for (int k = 0; k < n; k++)
   {
      int off = 0;
      out[k] = (float2 *)device_field(s[k], F32, &off);
      if (!out[k] || off)
      {
         fail(where, "the fields must be full single-precision device fields");
         return false;
      }
   }
   return true;
}

}  // namespace

bool nn_ready() { return N.nl > 0 && lib().nn_prec; }

// phi += NN(r) for the preconditioner (device fields of the lattice): the
// activations between the layers are kept in BF16, the values the channel
// mixing rounds them to anyway, and the last layer adds its output to phi
bool nn_correct(const float2 *r, float2 *phi)
{
   if (!N.nl || !ensure_fields()) return false;
   pack();
   Lib &L = lib();
   if (N.u16_epoch != L.u_epoch)
   {
      // the transports of the preconditioner read the links in FP16 (half
      // the bytes of the layer's dominant stream; the BF16 channel mixing
      // rounds more coarsely)
      const size_t nu = units_for(L.g.V / 2, 72);
      k_ptc_link16<<<(unsigned)((nu + 255) / 256), 256, 0, L.stream>>>(L.d_u, N.d_u16, nu);
      N.u16_epoch = L.u_epoch;
   }
   const void *x = r;
   for (int l = 0; l < N.nl; l++)
   {
      const bool first = l == 0, last = l == N.nl - 1;
      const int mode = last ? (first ? 3 : 2) : (first ? 0 : 1);
      void *y = last ? (void *)phi : (void *)N.d_inf[l & 1];
      OQH_PTC_DISPATCH(launch_inf, N.w[l], N.w[l + 1], x, y, N.d_pf[l], mode);
      x = y;
   }
   return true;
}

float2 *nn_work() { return ensure_fields() ? N.d_tmp : nullptr; }

void release_nn() { free_net(); }

}  // namespace oqh

using namespace oqh;

extern "C" int oqh_nn_create(int nlayer, const int *widths)
{
   free_net();
   if (nlayer < 1 || nlayer > NLMAX || widths[0] != 4 || widths[nlayer] != 4)
   {
      fail("oqh_nn_create", "1 to 8 layers with 4 input and 4 output channels are supported");
      return 1;
   }
   for (int l = 0; l <= nlayer; l++)
      if (!valid_width(widths[l]))
      {
         fail("oqh_nn_create", "channel counts must be 4, 8, 12 or 16");
         return 1;
      }
   N.nl = nlayer;
   for (int l = 0; l <= nlayer; l++) N.w[l] = widths[l];
   for (int l = 0; l < nlayer; l++) N.off[l + 1] = N.off[l] + (size_t)NDIR * N.w[l] * N.w[l + 1];
   size_t nw = nweights();
   N.d_w = (float2 *)dev_alloc(nw * sizeof(float2));
   N.d_g = (float2 *)dev_alloc(nw * sizeof(float2));
   N.d_m = (float *)dev_alloc(2 * nw * sizeof(float));
   N.d_v = (float *)dev_alloc(2 * nw * sizeof(float));
   for (int l = 0; l < nlayer; l++)
   {
      N.d_pf[l] = (__bf16 *)dev_alloc((size_t)NDIR * 32 * kpad(N.w[l]) * sizeof(__bf16));
      N.d_pa[l] = (__bf16 *)dev_alloc((size_t)NDIR * 32 * kpad(N.w[l + 1]) * sizeof(__bf16));
   }
   N.d_lpart = (double *)dev_alloc(2 * 256 * sizeof(double));
   HIP_OK(hipHostMalloc((void **)&N.h_lpart, 2 * 256 * sizeof(double), 0));
   HIP_OK(hipMemset(N.d_w, 0, nw * sizeof(float2)));
   HIP_OK(hipMemset(N.d_m, 0, 2 * nw * sizeof(float)));
   HIP_OK(hipMemset(N.d_v, 0, 2 * nw * sizeof(float)));
   return 0;
}

extern "C" int oqh_nn_nweights() { return (int)nweights(); }

extern "C" void oqh_nn_set_weights(const double *w)
{
   if (!N.nl)
   {
      fail("oqh_nn_set_weights", "the network is not created");
      return;
   }
   size_t nw = nweights();
   std::vector<float2> h(nw);
   for (size_t k = 0; k < nw; k++) h[k] = make_float2((float)w[2 * k], (float)w[2 * k + 1]);
   HIP_OK(hipMemcpy(N.d_w, h.data(), nw * sizeof(float2), hipMemcpyHostToDevice));
   HIP_OK(hipMemset(N.d_m, 0, 2 * nw * sizeof(float)));
   HIP_OK(hipMemset(N.d_v, 0, 2 * nw * sizeof(float)));
   N.adam_t = 0;
   N.packed = false;
}

extern "C" void oqh_nn_get_weights(double *w)
{
   if (!N.nl)
   {
      fail("oqh_nn_get_weights", "the network is not created");
      return;
   }
   size_t nw = nweights();
   std::vector<float2> h(nw);
   HIP_OK(hipStreamSynchronize(lib().stream));
   HIP_OK(hipMemcpy(h.data(), N.d_w, nw * sizeof(float2), hipMemcpyDeviceToHost));
   for (size_t k = 0; k < nw; k++)
   {
      w[2 * k] = h[k].x;
      w[2 * k + 1] = h[k].y;
   }
}

extern "C" void oqh_nn_apply(spinor *in, spinor *out)
{
   const char *where = "oqh_nn_apply";
   if (!require_net(where)) return;
   Lib &L = lib();
   int V = L.g.V;
   Res a = resolve(in, F32, V, 0, true);
   Res b = resolve(out, F32, V, 1, false);
   if (!a.dev || !b.dev) return;
   if (a.off || b.off || a.dev == b.dev)
   {
      fail(where, "the fields must be distinct full fields");
      return;
   }
   forward((const float2 *)a.dev, (float2 *)b.dev);
   finish(a, F32, V, false);
   finish(b, F32, V, true);
}

// phi += NN(r) as the preconditioner applies it (BF16 activations between
// the layers); full single-precision device fields
extern "C" void oqh_nn_correct(spinor *r, spinor *phi)
{
   const char *where = "oqh_nn_correct";
   if (!require_net(where)) return;
   int o0 = 0, o1 = 0;
   float2 *a = (float2 *)device_field(r, F32, &o0), *b = (float2 *)device_field(phi, F32, &o1);
   if (!a || !b || o0 || o1 || a == b)
   {
      fail(where, "the fields must be distinct full single-precision device fields");
      return;
   }
   nn_correct(a, b);
   HIP_OK(hipStreamSynchronize(lib().stream));
}

extern "C" double oqh_nn_grad(int n, spinor **r, spinor **e, double *dw)
{
   const char *where = "oqh_nn_grad";
   if (!require_net(where)) return 0.0;
   std::vector<float2 *> rr, ee;
   if (!fields_of(n, r, rr, where) || !fields_of(n, e, ee, where)) return 0.0;
   double loss = loss_and_grad(n, rr.data(), ee.data(), N.d_tmp);
   size_t nw = nweights();
   std::vector<float2> h(nw);
   HIP_OK(hipStreamSynchronize(lib().stream));
   HIP_OK(hipMemcpy(h.data(), N.d_g, nw * sizeof(float2), hipMemcpyDeviceToHost));
   for (size_t k = 0; k < nw; k++)
   {
      dw[2 * k] = h[k].x;
      dw[2 * k + 1] = h[k].y;
   }
   return loss;
}

extern "C" double oqh_nn_train_step(int n, spinor **r, spinor **e, double lr)
{
   const char *where = "oqh_nn_train_step";
   if (!require_net(where)) return 0.0;
   std::vector<float2 *> rr, ee;
   if (!fields_of(n, r, rr, where) || !fields_of(n, e, ee, where)) return 0.0;
   Lib &L = lib();
   double loss = loss_and_grad(n, rr.data(), ee.data(), N.d_tmp);
   const float b1 = 0.9f, b2 = 0.999f;
   N.adam_t++;
   const float c1 = 1.0f / (1.0f - powf(b1, (float)N.adam_t)), c2 = 1.0f / (1.0f - powf(b2, (float)N.adam_t));
   const int nr = (int)(2 * nweights());
   k_adam<<<(nr + 255) / 256, 256, 0, L.stream>>>((float *)N.d_w, (const float *)N.d_g, N.d_m, N.d_v, nr, (float)lr, b1,
                                                  b2, 1e-8f, c1, c2);
   N.packed = false;
   HIP_OK(hipStreamSynchronize(L.stream));
   return loss;
}

extern "C" void oqh_set_nn_prec(int on)
{
   if (on != 0 && on != 1)
   {
      fail("oqh_set_nn_prec", "the argument must be 0 or 1");
      return;
   }
   if (on && !N.nl)
   {
      fail("oqh_set_nn_prec", "the network is not created");
      return;
   }
   lib().nn_prec = on;
}

extern "C" void oqh_nn_free() { free_net(); lib().nn_prec = 0; }
