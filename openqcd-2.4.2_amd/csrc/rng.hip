// rng.hip -- the reference's random number streams on the device and the
// random fields drawn from them.
//
// ranlxd (random/ranlxd.c, ranlux_common.c): four interleaved 48-bit
// subtract-with-borrow lanes, 12 words each plus the carry, advanced by
// p = 202 (level 1) or 397 (level 2) steps per refill of 48 doubles.
// start_ranlux (random/ranlux.c:136-171) gives reference thread k the seed
// seed + k (2^31 - 1) / NTHREAD.  random_ud (uflds/uflds.c:159-222) and
// random_sd (sflds/sflds.c:1161-1200) let each thread draw, in order, the
// values of its own points; here one lane runs the stream of one reference
// thread, so the fields are the reference's for the same seed and thread
// layout (up to the last bit of the device sin/cos/log in the Gaussian
// transform).  The stream state of a lane lives in LDS during a kernel and
// in a device array between kernels.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "dirac_dev.h"
#include "oqh_internal.h"

namespace oqh {

namespace {

constexpr long long RLX_BASE = 0x1000000000000LL;
constexpr long long RLX_MASK = 0xffffffffffffLL;
constexpr int RWG = 64;  // lanes (streams) per work-group

struct RlxState
{
   long long w[13][4];  // 12 words per lane plus the carries (row 12)
   int ir, pr, is, is_old;
   double rd[48];
};

RlxState *d_streams = nullptr;
int n_streams = 0;

// seed expansion of ranlux_common.c: 31-bit shift register, 24 bits per
// word, the words of the four lanes complemented by the flag pattern
void seed_state(RlxState &s, int seed, int level)
{
   int bits[31];
   for (int k = 0; k < 31; k++)
   {
      bits[k] = seed & 1;
      seed >>= 1;
   }
   int p = 0, q = 18;
   for (int lane = 0; lane < 4; lane++)
      for (int k = 0; k < 24; k++)
      {
         int word = 0;
         for (int l = 0; l < 24; l++)
         {
            word = 2 * word + bits[p];
            bits[p] = (bits[p] + bits[q]) & 1;
            p = (p + 1) % 31;
            q = (q + 1) % 31;
         }
         // ranlxd: every word but the lane's own quarter is complemented
         if ((k % 4) != lane) word = 0xffffff - word;
         if (k & 1)
            s.w[k / 2][lane] += (long long)word << 24;
         else
            s.w[k / 2][lane] = word;
      }
   for (int lane = 0; lane < 4; lane++) s.w[12][lane] = 0;
   s.ir = 0;
   s.pr = (level == 1) ? 202 : 397;
   s.is = 47;
   s.is_old = 0;
   memset(s.rd, 0, sizeof(s.rd));
}

// one lane's stream held in LDS
struct LaneRng
{
   long long (*w)[4];
   double *rd;
   int ir, pr, is, is_old;

   __device__ void refill()
   {
      int i = ir, j = (ir >= 5) ? ir - 5 : ir + 7;
      for (int k = 0; k < pr; k++)
      {
#pragma unroll
         for (int l = 0; l < 4; l++)
         {
            long long d = w[j][l] - w[i][l] - w[12][l];
            w[12][l] = d < 0;
            w[i][l] = (d + RLX_BASE) & RLX_MASK;
         }
         i = (i == 11) ? 0 : i + 1;
         j = (j == 11) ? 0 : j + 1;
      }
      ir = (ir + pr) % 12;
      const double unit = 3.552713678800501e-15;  // 2^-48
      for (int k = 0; k < 12; k++)
#pragma unroll
         for (int l = 0; l < 4; l++) rd[4 * k + l] = (double)w[k][l] * unit;
   }

   __device__ double next()
   {
      is = (is == 47) ? 0 : is + 1;
      if (is == is_old)
      {
         refill();
         is = 4 * ir;
         is_old = is;
      }
      return rd[is];
   }

   // gauss_dble (random/gauss.c): pairs of uniforms to pairs of normal
   // deviates with density ~ exp(-x^2); n even
   __device__ void gauss(double *r, int n)
   {
      for (int k = 0; k < n; k++) r[k] = next();
      for (int k = 0; k < n; k += 2)
      {
         const double rho = sqrt(-log(1.0 - r[k]));
         double sn, cs;
         sincos(6.2831853071795865 * (r[k + 1] - 0.5), &sn, &cs);
         r[k] = rho * sn;
         r[k + 1] = rho * cs;
      }
   }
};

__device__ inline void load_lane(const RlxState *g, long long (*w)[4], double *rd, LaneRng &R)
{
   for (int k = 0; k < 13; k++)
      for (int l = 0; l < 4; l++) w[k][l] = g->w[k][l];
   for (int k = 0; k < 48; k++) rd[k] = g->rd[k];
   R.w = w;
   R.rd = rd;
   R.ir = g->ir;
   R.pr = g->pr;
   R.is = g->is;
   R.is_old = g->is_old;
}

__device__ inline void store_lane(RlxState *g, const LaneRng &R)
{
   for (int k = 0; k < 13; k++)
      for (int l = 0; l < 4; l++) g->w[k][l] = R.w[k][l];
   for (int k = 0; k < 48; k++) g->rd[k] = R.rd[k];
   g->ir = R.ir;
   g->is = R.is;
   g->is_old = R.is_old;
}

// conj(a x b) of two colour vectors (6 reals each)
__device__ inline void cross_conj6(const double *a, const double *b, double *c)
{
   for (int i = 0; i < 3; i++)
   {
      const int j = (i + 1) % 3, k = (i + 2) % 3;
      const double re = (a[2 * j] * b[2 * k] - a[2 * j + 1] * b[2 * k + 1]) -
                        (a[2 * k] * b[2 * j] - a[2 * k + 1] * b[2 * j + 1]);
      const double im = (a[2 * j] * b[2 * k + 1] + a[2 * j + 1] * b[2 * k]) -
                        (a[2 * k] * b[2 * j + 1] + a[2 * k + 1] * b[2 * j]);
      c[2 * i] = re;
      c[2 * i + 1] = -im;
   }
}

// random unit colour vector (norm^2 of the Gaussian draw above 0.1)
__device__ inline void unit_vec3(LaneRng &R, double *v)
{
   double n2 = 0.0;
   while (n2 <= 0.1)
   {
      R.gauss(v, 6);
      n2 = 0.0;
      for (int i = 0; i < 6; i++) n2 += v[i] * v[i];
   }
   const double f = 1.0 / sqrt(n2);
   for (int i = 0; i < 6; i++) v[i] *= f;
}

// random_su3_dble (su3fcts/random_su3.c): rows r0, r1 = conj(r2 x r0) with
// r2 = conj(r0 x r1') normalised
__device__ inline void random_su3(LaneRng &R, double *u)
{
   double *r0 = u, *r1 = u + 6, *r2 = u + 12;
   unit_vec3(R, r0);
   double n2 = 0.0;
   while (n2 <= 0.1)
   {
      unit_vec3(R, r1);
      cross_conj6(r0, r1, r2);
      n2 = 0.0;
      for (int i = 0; i < 6; i++) n2 += r2[i] * r2[i];
   }
   const double f = 1.0 / sqrt(n2);
   for (int i = 0; i < 6; i++) r2[i] *= f;
   cross_conj6(r2, r0, r1);
}

// random_ud: lane k draws the 8 links of the odd points of reference thread
// k in order (boundary links of open bc are not drawn and are set to zero)
__global__ __launch_bounds__(RWG) void k_random_ud(RlxState *streams, int nthr, double2 *ud, const int *x0, int h,
                                                   int vt2, int N0, int bc)
{
   __shared__ long long W[RWG][13][4];
   __shared__ double RD[RWG][48];
   const int k = blockIdx.x * RWG + threadIdx.x;
   if (k >= nthr) return;
   LaneRng R;
   load_lane(streams + k, W[threadIdx.x], RD[threadIdx.x], R);
   double u[18];
   for (int j = k * vt2; j < (k + 1) * vt2; j++)
   {
      const int t = x0[h + j];
      for (int l = 0; l < 8; l++)
      {
         const bool skip = bc == 0 && ((t == 0 && l == 1) || (t == N0 - 1 && l == 0));
         if (!skip)
            random_su3(R, u);
         else
            for (int c = 0; c < 18; c++) u[c] = 0.0;
         double2 *p = ud + (size_t)(j >> 6) * (72 * SB) + (j & (SB - 1)) + (size_t)(9 * l) * SB;
         for (int c = 0; c < 9; c++) p[c * SB] = make_double2(u[2 * c], u[2 * c + 1]);
      }
   }
   store_lane(streams + k, R);
}

// random_sd: lane k (stream k0 + k) fills the points [x0 + k chunk,
// x0 + (k+1) chunk) with sigma times Gaussian deviates (24 per point, in the
// spinor's component order)
template <typename C>
__global__ __launch_bounds__(RWG) void k_random_sf(RlxState *streams, int k0, int nthr, C *sd, int x0, int chunk,
                                                   double sigma)
{
   __shared__ long long W[RWG][13][4];
   __shared__ double RD[RWG][48];
   const int k = blockIdx.x * RWG + threadIdx.x;
   if (k >= nthr) return;
   LaneRng R;
   load_lane(streams + k0 + k, W[threadIdx.x], RD[threadIdx.x], R);
   double r[24];
   for (int x = x0 + k * chunk; x < x0 + (k + 1) * chunk; x++)
   {
      R.gauss(r, 24);
      C *p = sd + (size_t)(x >> 6) * (12 * SB) + (x & (SB - 1));
      for (int c = 0; c < 12; c++) p[c * SB] = cmk<C>(sigma * r[2 * c], sigma * r[2 * c + 1]);
   }
   store_lane(streams + k0 + k, R);
}

// random source of main/ms4.c:383-406: lane k walks the ipt indices of
// reference thread k in order and draws a Gaussian spinor (random_sd(1,0,..)
// with its stream) at the points of time slice t0, zero elsewhere
__global__ __launch_bounds__(RWG) void k_random_source(RlxState *streams, int nthr, double2 *sd, const int *x0c,
                                                       int vt, int t0)
{
   __shared__ long long W[RWG][13][4];
   __shared__ double RD[RWG][48];
   const int k = blockIdx.x * RWG + threadIdx.x;
   if (k >= nthr) return;
   LaneRng R;
   load_lane(streams + k, W[threadIdx.x], RD[threadIdx.x], R);
   double r[24];
   for (int ix = k * vt; ix < (k + 1) * vt; ix++)
   {
      // thread k of ms4.c walks the ipt indices [k VOLUME_TRD, (k+1) VOLUME_TRD)
      double2 *p = sd + (size_t)(ix >> 6) * (12 * SB) + (ix & (SB - 1));
      if (x0c[ix] == t0)
      {
         R.gauss(r, 24);
         for (int c = 0; c < 12; c++) p[c * SB] = make_double2(r[2 * c], r[2 * c + 1]);
      }
      else
         for (int c = 0; c < 12; c++) p[c * SB] = make_double2(0.0, 0.0);
   }
   store_lane(streams + k, R);
}

// n numbers of stream 0 (the master thread's ranlxd)
__global__ void k_ranlxd(RlxState *streams, double *out, int n)
{
   __shared__ long long W[13][4];
   __shared__ double RD[48];
   if (threadIdx.x != 0) return;
   LaneRng R;
   load_lane(streams, W, RD, R);
   for (int k = 0; k < n; k++) out[k] = R.next();
   store_lane(streams, R);
}

bool streams_ready(const char *where)
{
   Lib &L = lib();
   if (!require_geometry(where)) return false;
   if (!d_streams || n_streams != L.g.nthread)
   {
      fail(where, "the random number generator is not initialized (start_ranlux)");
      return false;
   }
   return true;
}

}  // namespace

// random start fields of the deflation subspace generation
// Philox4x32-10 (J. K. Salmon, M. A. Moraes, R. O. Dror, D. E. Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC11, 2011): ten rounds of
// two 32x32->64-bit products with the key bumped by the Weyl constants
__device__ inline uint4 philox4x32_10(uint4 c, uint32_t k0, uint32_t k1)
{
#pragma unroll
   for (int r = 0; r < 10; r++)
   {
      const uint64_t p0 = (uint64_t)0xD2511F53u * c.x, p1 = (uint64_t)0xCD9E8D57u * c.z;
      c = make_uint4((uint32_t)(p1 >> 32) ^ c.y ^ k0, (uint32_t)p1, (uint32_t)(p0 >> 32) ^ c.w ^ k1, (uint32_t)p0);
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
   }
   return c;
}

// start fields of the subspace generation: component 4 j + i of point x
// (ipt index) is word i of Philox4x32-10 at counter (x, j, 0, 0), j < 6, as
// (w >> 8) 2^-24 - 1/2; the key is two ranlxd numbers of stream 0 scaled
// by 2^32 (key[0], key[1] on the device)
__global__ __launch_bounds__(256) void k_random_start(float2 *f, int V, const double *key)
{
   const int x = blockIdx.x * 256 + threadIdx.x;
   if (x >= V) return;
   const uint32_t k0 = (uint32_t)(key[0] * 4294967296.0), k1 = (uint32_t)(key[1] * 4294967296.0);
   float2 *p = f + (size_t)(x >> 6) * (12 * SB) + (x & (SB - 1));
   const float s = 1.0f / 16777216.0f;
#pragma unroll
   for (int j = 0; j < 6; j++)
   {
      const uint4 w = philox4x32_10(make_uint4((uint32_t)x, (uint32_t)j, 0u, 0u), k0, k1);
      p[(2 * j) * SB] = make_float2((float)(w.x >> 8) * s - 0.5f, (float)(w.y >> 8) * s - 0.5f);
      p[(2 * j + 1) * SB] = make_float2((float)(w.z >> 8) * s - 0.5f, (float)(w.w >> 8) * s - 0.5f);
   }
}

double *d_key = nullptr;

void random_uniform_s(float2 *f)
{
   Lib &L = lib();
   if (!streams_ready("dfl_modes [dfl_modes.c]") || !f) return;
   if (!d_key) d_key = (double *)dev_alloc(sizeof(double) * 2);
   if (!d_key) return;
   k_ranlxd<<<1, 64, 0, L.stream>>>(d_streams, d_key, 2);
   k_random_start<<<(L.g.V + 255) / 256, 256, 0, L.stream>>>(f, L.g.V, d_key);
}

void release_rng()
{
   dev_free(d_streams);
   d_streams = nullptr;
   dev_free(d_key);
   d_key = nullptr;
   n_streams = 0;
}

}  // namespace oqh

using namespace oqh;

extern "C" void start_ranlux(int level, int seed)
{
   const char *where = "start_ranlux [ranlux.c]";
   Lib &L = lib();
   if (!require_geometry(where)) return;
   if (level < 0 || level > 1 || seed < 1)
   {
      fail(where, "Parameters are out of range");
      return;
   }
   const int nthr = L.g.nthread;
   const int max_seed = 2147483647 / nthr;
   if (seed >= max_seed)
   {
      fail(where, "Parameters are out of range");
      return;
   }
   std::vector<RlxState> h(nthr);
   for (int k = 0; k < nthr; k++) seed_state(h[k], seed + k * max_seed, level + 1);
   if (n_streams != nthr)
   {
      release_rng();
      d_streams = (RlxState *)dev_alloc(sizeof(RlxState) * nthr);
      if (!d_streams) return;
      n_streams = nthr;
   }
   HIP_OK(hipMemcpy(d_streams, h.data(), sizeof(RlxState) * nthr, hipMemcpyHostToDevice));
}

extern "C" void random_ud(void)
{
   const char *where = "random_ud [uflds.c]";
   Lib &L = lib();
   if (!streams_ready(where)) return;
   const int h = L.g.V / 2, nthr = L.g.nthread;
   k_random_ud<<<(nthr + RWG - 1) / RWG, RWG, 0, L.stream>>>(d_streams, nthr, L.d_ud, L.d_x0, h, L.g.VT / 2,
                                                            L.g.L[0], L.bc);
   HIP_OK(hipStreamSynchronize(L.stream));
   L.ud_set = true;
   L.ud_phase = false;
   L.ud_version++;
   L.u_valid = false;
   L.swd_valid = L.sw_valid = false;
}

// random_sd (sflds/sflds.c:1254-1268): icom & 2: every
// reference thread k fills vol points from s + k vol with its own stream;
// otherwise the calling (master) thread fills vol points from s with stream 0
template <typename C>
static void random_field(int vol, int icom, void *s, double sigma, int prec, const char *where)
{
   Lib &L = lib();
   if (!streams_ready(where)) return;
   const int nthr = (icom & 0x2) ? L.g.nthread : 1;
   const int n = vol * nthr;
   if (vol <= 0)
   {
      fail(where, "Parameters are out of range");
      return;
   }
   Res r = resolve(s, prec, n, 0, false);
   if (!r.dev) return;
   k_random_sf<C><<<(nthr + RWG - 1) / RWG, RWG, 0, L.stream>>>(d_streams, 0, nthr, (C *)r.dev, r.off, vol, sigma);
   finish(r, prec, n, true);
}

extern "C" void random_sd(int vol, int icom, spinor_dble *sd, double sigma)
{
   random_field<double2>(vol, icom, sd, sigma, F64, "random_sd [sflds.c]");
}

// the random source of main/ms4.c:383-406 on time slice t0 (full field)
// one start field of the subspace generation (dfl_modes), as test access
extern "C" void oqh_random_start_field(spinor *s)
{
   const char *where = "oqh_random_start_field";
   Lib &L = lib();
   if (!streams_ready(where)) return;
   Res r = resolve(s, F32, L.g.V, 0, false);
   if (!r.dev) return;
   if (r.off)
   {
      fail(where, "the field must be a full field");
      return;
   }
   random_uniform_s((float2 *)r.dev);
   finish(r, F32, L.g.V, true);
}

extern "C" void oqh_random_source(int t0, spinor_dble *sd)
{
   const char *where = "oqh_random_source";
   Lib &L = lib();
   if (!streams_ready(where)) return;
   if (t0 < 0 || t0 >= L.g.L[0])
   {
      fail(where, "the time slice is out of range");
      return;
   }
   Res r = resolve(sd, F64, L.g.V, 0, false);
   if (!r.dev) return;
   if (r.off)
   {
      fail(where, "the field must be a full field");
      return;
   }
   const int nthr = L.g.nthread;
   k_random_source<<<(nthr + RWG - 1) / RWG, RWG, 0, L.stream>>>(d_streams, nthr, (double2 *)r.dev, L.d_x0, L.g.VT,
                                                                t0);
   finish(r, F64, L.g.V, true);
}

extern "C" void ranlxd(double *r, int n)
{
   const char *where = "ranlxd [ranlxd.c]";
   Lib &L = lib();
   if (!streams_ready(where) || n <= 0) return;
   double *d = (double *)tmp_buffer(sizeof(double) * n);
   k_ranlxd<<<1, 64, 0, L.stream>>>(d_streams, d, n);
   HIP_OK(hipMemcpyAsync(r, d, sizeof(double) * n, hipMemcpyDeviceToHost, L.stream));
   HIP_OK(hipStreamSynchronize(L.stream));
}
