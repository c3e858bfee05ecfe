// rng.hip -- the reference's random number streams on the device and the
// random fields drawn from them.
//
// ranlxd (random/ranlxd.c, ranlux_common.c): four interleaved 48-bit
// subtract-with-borrow lanes, 12 words each plus the carry, advanced by
// p = 202 (level 1) or 397 (level 2) steps per refill of 48 doubles.
// start_ranlux (random/ranlux.c:136-171) gives reference thread k the seed
// seed + k (2^31 - 1) / NTHREAD.  random_ud (uflds/uflds.c:159-222) and
// random_sd (sflds/sflds.c:1161-1200) let each thread draw, in order, the
// values of its own points; the device runs the stream of every reference
// thread, so the fields are the reference's for the same seed and thread
// layout (up to the last bit of the device sin/cos/log in the Gaussian
// transform).  The stream states live in a device array between kernels.
// Two ways to run the streams give the same numbers and leave the same
// states: by default every refill of every stream is reached by a jump and
// the draws are assigned to the points in parallel (all streams of one
// call at once, whatever the thread layout); with OQH_RNG_SERIAL=1 one lane
// walks each stream, its state in LDS (also the fallback when a chain of
// random_ud draws does not fit the parallel bookkeeping).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "dirac_dev.h"
#include "oqh_internal.h"
#include "rlx_jump.h"

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

// The arithmetic of the random fields is not contracted to fused
// multiply-adds, so that every kernel that draws a field rounds as the
// reference's scalar code does and the per-stream and parallel paths agree
// bit for bit.

// gauss_dble's transform (random/gauss.c): a pair of uniforms to a pair of
// normal deviates with density ~ exp(-x^2)
__device__ inline void gauss_pair(double u0, double u1, double &g0, double &g1)
{
#pragma clang fp contract(off)
   const double rho = sqrt(-log(1.0 - u0));
   double sn, cs;
   sincos(6.2831853071795865 * (u1 - 0.5), &sn, &cs);
   g0 = rho * sn;
   g1 = rho * cs;
}

// square norm of a colour vector (6 reals), summed in order
__device__ inline double norm6(const double *v)
{
#pragma clang fp contract(off)
   double n2 = 0.0;
   for (int i = 0; i < 6; i++) n2 += v[i] * v[i];
   return n2;
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

   // gauss_dble (random/gauss.c): n uniforms, then pairs of them to pairs of
   // normal deviates; n even
   __device__ void gauss(double *r, int n)
   {
      for (int k = 0; k < n; k++) r[k] = next();
      for (int k = 0; k < n; k += 2) gauss_pair(r[k], r[k + 1], r[k], r[k + 1]);
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
#pragma clang fp contract(off)
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
      n2 = norm6(v);
   }
   const double f = 1.0 / sqrt(n2);
   for (int i = 0; i < 6; i++) v[i] *= f;
}

// the last two rows of random_su3 once r2 = conj(r0 x r1') has passed the
// test (n2 its square norm): r2 normalised, r1 = conj(r2 x r0)
__device__ inline void su3_rows(double *u, double n2)
{
#pragma clang fp contract(off)
   double *r0 = u, *r1 = u + 6, *r2 = u + 12;
   const double f = 1.0 / sqrt(n2);
   for (int i = 0; i < 6; i++) r2[i] *= f;
   cross_conj6(r2, r0, r1);
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
      n2 = norm6(r2);
   }
   su3_rows(u, n2);
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

// ---------------------------------------------------------------------------
// parallel generation.  The numbers a stream hands out during one call are
// made by jumping each lane of each stream to the refills it needs
// (rlx_jump.h), so that all refills of all streams are computed at once; the
// draws are then assigned to the points by their counts.  These are fixed
// for random_sd and the random source.  For random_ud the acceptance tests of
// random_su3 make the count of a link variable: the uniforms are cut into
// groups of six (one unit vector each), every group g gets the groups a0(g),
// a1(g) of the link that would start there, and the link starts follow as
// the chain g -> a1(g) + 1 from group 0, resolved in chunks of groups.  The
// values are those of the per-stream walk, number for number.
// ---------------------------------------------------------------------------

struct JumpTable
{
   uint32_t p[31][rlx::NR];  // a^(pr 2^i) mod m
};

constexpr int PWG = 256;    // work-group of the parallel kernels
constexpr int NO_LINK = -1; // a1(g) when the buffer ends before the link does
constexpr int X_END = -1;   // chunk exit: a link without end
constexpr int X_FAR = -2;   // chunk exit: beyond the entry window

// numbers left in the current refill of a stream
__device__ inline int tail_count(const RlxState &S) { return (S.is_old - S.is - 1 + 96) % 48; }

// words (chronological) and borrow of lane l
__device__ inline void lane_words(const RlxState &S, int l, uint64_t *x, int *c)
{
   for (int m = 0; m < 12; m++) x[m] = (uint64_t)S.w[(S.ir + m) % 12][l];
   *c = (int)S.w[12][l];
}

// z <- z a^(pr j)
__device__ inline void jump_refills(uint32_t *z, long long j, const JumpTable *tab)
{
   for (int i = 0; j; i++, j >>= 1)
      if (j & 1) rlx::mulmod(z, tab->p[i], z);
}

// the first numbers of stream s: those left in its current refill
__global__ void k_rlx_tail(const RlxState *streams, int nstr, double *U, long long nu)
{
   const int s = blockIdx.x, k = threadIdx.x;
   if (s >= nstr || k >= 48) return;
   const RlxState &S = streams[s];
   if (k < tail_count(S) && k < nu) U[(size_t)s * nu + k] = S.rd[(S.is + 1 + k) % 48];
}

// the refills of the streams: U[s nu + n0 + 48 (j - 1) + 4 m + l] = word m
// (chronological) of lane l after refill j; thread (s, t, l) makes refills
// 1 + t K .. (t + 1) K by one jump and one product per further refill
__global__ __launch_bounds__(PWG) void k_rlx_produce(const RlxState *streams, int nstr, int nchunk, int K, double *U,
                                                     long long nu, const JumpTable *tab, int *bad)
{
   const long long id = (long long)blockIdx.x * PWG + threadIdx.x;
   if (id >= (long long)nstr * nchunk * 4) return;
   const int l = (int)(id & 3);
   const long long q = id >> 2;
   const int t = (int)(q % nchunk), s = (int)(q / nchunk);
   const RlxState &S = streams[s];
   const long long n0 = tail_count(S), j0 = 1 + (long long)t * K;
   if (n0 + 48 * (j0 - 1) >= nu) return;
   double *u = U + (size_t)s * nu;
   uint64_t x[12];
   int c;
   lane_words(S, l, x, &c);
   uint32_t z[rlx::NR];
   rlx::state_to_z(x, c, z);
   jump_refills(z, j0, tab);
   const double unit = 3.552713678800501e-15;  // 2^-48
   for (long long j = j0; j < j0 + K; j++)
   {
      const long long base = n0 + 48 * (j - 1);
      if (base >= nu) break;
      if (j > j0) rlx::mulmod(z, tab->p[0], z);
      if (!rlx::z_to_state(z, x, &c))
      {
         *bad = 1;
         return;
      }
      for (int m = 0; m < 12; m++)
      {
         const long long k = base + 4 * m + l;
         if (k < nu) u[k] = (double)x[m] * unit;
      }
   }
}

// stream s after cons[s] numbers: the refill that holds the last one, its
// position and the words and borrows after it
__global__ void k_rlx_finish(RlxState *streams, int nstr, const long long *cons, long long nu,
                             const JumpTable *tab, int *bad)
{
   const int s = blockIdx.x * 64 + threadIdx.x;
   if (s >= nstr) return;
   RlxState &S = streams[s];
   const long long n = cons[s];
   const int n0 = tail_count(S);
   if (n < 0 || n > nu)
   {
      *bad = 1;
      return;
   }
   if (n <= n0)
   {
      S.is = (int)((S.is + n) % 48);
      return;
   }
   const long long jn = (n - n0 + 47) / 48;
   const int clast = (int)(n - n0 - 48 * (jn - 1));
   const int irn = (int)((S.ir + jn * S.pr) % 12);
   uint64_t nw[4][12];
   int nc[4];
   for (int l = 0; l < 4; l++)
   {
      uint64_t x[12];
      int c;
      lane_words(S, l, x, &c);
      uint32_t z[rlx::NR];
      rlx::state_to_z(x, c, z);
      jump_refills(z, jn, tab);
      if (!rlx::z_to_state(z, nw[l], &nc[l]))
      {
         *bad = 1;
         return;
      }
   }
   const double unit = 3.552713678800501e-15;
   for (int l = 0; l < 4; l++)
   {
      for (int m = 0; m < 12; m++) S.w[(irn + m) % 12][l] = (long long)nw[l][m];
      S.w[12][l] = nc[l];
   }
   for (int k = 0; k < 12; k++)
      for (int l = 0; l < 4; l++) S.rd[4 * k + l] = (double)S.w[k][l] * unit;
   S.ir = irn;
   S.is_old = 4 * irn;
   S.is = (S.is_old + clast - 1) % 48;
}

// random_su3's unit vectors: group gg (six uniforms) to its Gaussian
// vector, the acceptance |v|^2 > 0.1 and, if accepted, the unit vector
__global__ __launch_bounds__(PWG) void k_rlx_unit6(double *U, long long ng, unsigned char *acc)
{
   const long long gg = (long long)blockIdx.x * PWG + threadIdx.x;
   if (gg >= ng) return;
   double v[6];
   for (int k = 0; k < 6; k++) v[k] = U[6 * gg + k];
   for (int k = 0; k < 6; k += 2) gauss_pair(v[k], v[k + 1], v[k], v[k + 1]);
   const double n2 = norm6(v);
   const bool ok = n2 > 0.1;
   acc[gg] = ok ? 1 : 0;
   if (ok)
   {
      const double f = 1.0 / sqrt(n2);
      for (int i = 0; i < 6; i++) v[i] *= f;
   }
   for (int k = 0; k < 6; k++) U[6 * gg + k] = v[k];
}

// the link that would start at group g of stream s: r0 from the first
// accepted group a0 >= g, r1' from the first accepted group a1 > a0 for
// which conj(r0 x r1') has square norm above 0.1
__global__ __launch_bounds__(PWG) void k_rlx_links(const double *U, const unsigned char *acc, int nstr, int G,
                                                   int *A0, int *A1)
{
   const long long gg = (long long)blockIdx.x * PWG + threadIdx.x;
   if (gg >= (long long)nstr * G) return;
   const int s = (int)(gg / G), g = (int)(gg % G);
   const unsigned char *ac = acc + (size_t)s * G;
   const double *V = U + (size_t)6 * s * G;
   int a0 = g;
   while (a0 < G && !ac[a0]) a0++;
   int a1 = NO_LINK;
   if (a0 < G)
   {
      double r0[6], r1[6], r2[6];
      for (int i = 0; i < 6; i++) r0[i] = V[6 * (size_t)a0 + i];
      for (int b = a0 + 1; b < G; b++)
      {
         if (!ac[b]) continue;
         for (int i = 0; i < 6; i++) r1[i] = V[6 * (size_t)b + i];
         cross_conj6(r0, r1, r2);
         if (norm6(r2) > 0.1)
         {
            a1 = b;
            break;
         }
      }
   }
   A0[gg] = a0;
   A1[gg] = a1;
}

// chunk c of stream s entered at group c C + e: the number of link starts in
// the chunk and the entry into the next chunk (or X_END, X_FAR)
__global__ __launch_bounds__(PWG) void k_chain_walk(const int *A1, int nstr, int G, int C, int W, int *EX, int *CN)
{
   const int nch = G / C;
   const long long id = (long long)blockIdx.x * PWG + threadIdx.x;
   if (id >= (long long)nstr * nch * W) return;
   const int e = (int)(id % W), c = (int)((id / W) % nch), s = (int)(id / ((long long)W * nch));
   const int *a1 = A1 + (size_t)s * G;
   const int end = (c + 1) * C;
   int g = c * C + e, cnt = 0, ex;
   for (;;)
   {
      cnt++;
      const int b = a1[g];
      if (b < 0)
      {
         ex = X_END;
         break;
      }
      g = b + 1;
      if (g >= end)
      {
         ex = (g - end < W) ? g - end : X_FAR;
         break;
      }
   }
   EX[id] = ex;
   CN[id] = cnt;
}

// stream s: rank of the first link start and entry of every chunk, until
// the starts of ranks 0 .. NL[s] are found (CB = -1 for chunks not needed)
__global__ void k_chain_scan(const int *EX, const int *CN, int nstr, int nch, int W, const int *NL, int *CB, int *CE,
                             int *bad)
{
   const int s = blockIdx.x * 64 + threadIdx.x;
   if (s >= nstr) return;
   const int need = NL[s];
   int base = 0, e = 0, c = 0;
   bool ok = false;
   for (; c < nch; c++)
   {
      const size_t id = ((size_t)s * nch + c) * W + e;
      CB[(size_t)s * nch + c] = base;
      CE[(size_t)s * nch + c] = e;
      const int ex = EX[id], cn = CN[id];
      if (ex < 0)
      {
         // the chunk's last start has no tracked end: enough if its rank is
         // at least need
         ok = base + cn - 1 >= need;
         c++;
         break;
      }
      base += cn;
      e = ex;
      if (base > need)
      {
         ok = true;
         c++;
         break;
      }
   }
   for (; c < nch; c++) CB[(size_t)s * nch + c] = -1;
   if (!ok) *bad = 1;
}

// the link starts: S[s sst + k] = group of the link of rank k, k <= NL[s]
__global__ __launch_bounds__(PWG) void k_chain_fill(const int *A1, int nstr, int G, int C, const int *CB, const int *CE,
                                                    const int *NL, int *S, int sst)
{
   const int nch = G / C;
   const long long id = (long long)blockIdx.x * PWG + threadIdx.x;
   if (id >= (long long)nstr * nch) return;
   const int base = CB[id];
   if (base < 0) return;
   const int s = (int)(id / nch), c = (int)(id % nch);
   const int need = NL[s], end = (c + 1) * C;
   const int *a1 = A1 + (size_t)s * G;
   int g = c * C + CE[id];
   for (int k = base; k <= need; k++)
   {
      S[(size_t)s * sst + k] = g;
      const int b = a1[g];
      if (b < 0) break;
      g = b + 1;
      if (g >= end) break;
   }
}

// ranks of the first draw of each point of stream s (exclusive prefix sums
// of the draws per point) and their total NL[s]; MODE 0: the links of the
// odd point h + s npts + j that random_ud draws (8, less the open-bc
// boundary links); MODE 1: the points of time slice t0 (one draw each)
template <int MODE>
__global__ __launch_bounds__(PWG) void k_rlx_rank(const int *x0, int npts, int xbase, int N0, int bc, int t0,
                                                  int *rank, int *NL)
{
   __shared__ int sh[PWG];
   __shared__ int run;
   const int s = blockIdx.x;
   if (threadIdx.x == 0) run = 0;
   __syncthreads();
   for (int j0 = 0; j0 < npts; j0 += PWG)
   {
      const int j = j0 + threadIdx.x;
      int cnt = 0;
      if (j < npts)
      {
         const int t = x0[xbase + s * npts + j];
         if (MODE == 0)
            cnt = 8 - ((bc == 0 && t == 0) ? 1 : 0) - ((bc == 0 && t == N0 - 1) ? 1 : 0);
         else
            cnt = (t == t0) ? 1 : 0;
      }
      sh[threadIdx.x] = cnt;
      __syncthreads();
      for (int off = 1; off < PWG; off <<= 1)
      {
         const int v = (threadIdx.x >= off) ? sh[threadIdx.x - off] : 0;
         __syncthreads();
         sh[threadIdx.x] += v;
         __syncthreads();
      }
      const int incl = sh[threadIdx.x];
      if (j < npts) rank[(size_t)s * npts + j] = run + incl - cnt;
      __syncthreads();
      if (threadIdx.x == PWG - 1) run += incl;
      __syncthreads();
   }
   if (threadIdx.x == 0) NL[s] = run;
}

// random_ud from the resolved links: odd point j of stream s
__global__ __launch_bounds__(PWG) void k_rlx_ud(const double *U, const int *A0, const int *A1, const int *S, int sst,
                                                const int *rank, int nstr, int vt2, int G, double2 *ud, const int *x0,
                                                int h, int N0, int bc)
{
   const long long id = (long long)blockIdx.x * PWG + threadIdx.x;
   if (id >= (long long)nstr * vt2) return;
   const int s = (int)(id / vt2);
   const int j = (int)id;  // odd-point index: stream s holds [s vt2, (s+1) vt2)
   const int t = x0[h + j];
   const double *V = U + (size_t)6 * s * G;
   const int *a0 = A0 + (size_t)s * G, *a1 = A1 + (size_t)s * G, *st = S + (size_t)s * sst;
   int r = rank[id];
   double u[18];
   for (int l = 0; l < 8; l++)
   {
      const bool skip = bc == 0 && ((t == 0 && l == 1) || (t == N0 - 1 && l == 0));
      if (!skip)
      {
         const int g = st[r++];
         const double *p0 = V + 6 * (size_t)a0[g], *p1 = V + 6 * (size_t)a1[g];
         for (int i = 0; i < 6; i++)
         {
            u[i] = p0[i];
            u[6 + i] = p1[i];
         }
         cross_conj6(u, u + 6, u + 12);
         su3_rows(u, norm6(u + 12));
      }
      else
         for (int c = 0; c < 18; c++) u[c] = 0.0;
      double2 *p = ud + (size_t)(j >> 6) * (72 * SB) + (j & (SB - 1)) + (size_t)(9 * l) * SB;
      for (int c = 0; c < 9; c++) p[c * SB] = make_double2(u[2 * c], u[2 * c + 1]);
   }
}

// numbers drawn by random_ud: six per group up to the start of rank NL[s]
__global__ void k_ud_cons(const int *S, int sst, const int *NL, int nstr, long long *cons)
{
   const int s = blockIdx.x * 64 + threadIdx.x;
   if (s < nstr) cons[s] = 6LL * S[(size_t)s * sst + NL[s]];
}

// random_sd from the uniforms: point p of stream s is x0 + s vol + p, 24
// numbers each
template <typename C>
__global__ __launch_bounds__(PWG) void k_rlx_sf(const double *U, long long nu, int nstr, int vol, C *sd, int x0,
                                                double sigma)
{
   const long long id = (long long)blockIdx.x * PWG + threadIdx.x;
   if (id >= (long long)nstr * vol) return;
   const int s = (int)(id / vol), pnt = (int)(id % vol);
   const double *u = U + (size_t)s * nu + 24 * (size_t)pnt;
   double r[24];
   for (int k = 0; k < 24; k++) r[k] = u[k];
   for (int k = 0; k < 24; k += 2) gauss_pair(r[k], r[k + 1], r[k], r[k + 1]);
   const int x = x0 + (int)id;
   C *p = sd + (size_t)(x >> 6) * (12 * SB) + (x & (SB - 1));
   for (int c = 0; c < 12; c++) p[c * SB] = cmk<C>(sigma * r[2 * c], sigma * r[2 * c + 1]);
}

// the random source from the uniforms: ipt index ix of stream s = ix / vt
__global__ __launch_bounds__(PWG) void k_rlx_source(const double *U, long long nu, const int *rank, int V, int vt,
                                                    double2 *sd, const int *x0c, int t0)
{
   const int ix = blockIdx.x * PWG + threadIdx.x;
   if (ix >= V) return;
   const int s = ix / vt;
   double2 *p = sd + (size_t)(ix >> 6) * (12 * SB) + (ix & (SB - 1));
   if (x0c[ix] == t0)
   {
      const double *u = U + (size_t)s * nu + 24 * (size_t)rank[ix];
      double r[24];
      for (int k = 0; k < 24; k++) r[k] = u[k];
      for (int k = 0; k < 24; k += 2) gauss_pair(r[k], r[k + 1], r[k], r[k + 1]);
      for (int c = 0; c < 12; c++) p[c * SB] = make_double2(r[2 * c], r[2 * c + 1]);
   }
   else
      for (int c = 0; c < 12; c++) p[c * SB] = make_double2(0.0, 0.0);
}

// n numbers per stream for the source kernel
__global__ void k_src_cons(const int *NL, int nstr, long long *cons)
{
   const int s = blockIdx.x * 64 + threadIdx.x;
   if (s < nstr) cons[s] = 24LL * NL[s];
}

__global__ void k_fill_cons(long long *cons, int nstr, long long n)
{
   const int s = blockIdx.x * 64 + threadIdx.x;
   if (s < nstr) cons[s] = n;
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

JumpTable *d_jtab = nullptr;
int jtab_pr = 0;
int rlx_pr = 202;  // refill length of the streams (set by start_ranlux)

// OQH_RNG_SERIAL=1 selects the per-stream walk (one lane per stream)
bool rng_serial()
{
   const char *e = getenv("OQH_RNG_SERIAL");
   return e && e[0] == '1' && e[1] == 0;
}

// the products a^(pr 2^i) mod m of the current refill length
bool ensure_jtab()
{
   Lib &L = lib();
   if (d_jtab && jtab_pr == rlx_pr) return true;
   JumpTable h;
   uint32_t a[rlx::NR], p[rlx::NR];
   rlx::multiplier(a);
   for (int i = 0; i < rlx::NR; i++) p[i] = 0;
   p[0] = 1;
   for (int n = rlx_pr; n; n >>= 1)
   {
      if (n & 1) rlx::mulmod(p, a, p);
      rlx::mulmod(a, a, a);
   }
   memcpy(h.p[0], p, sizeof(p));
   for (int i = 1; i < 31; i++) rlx::mulmod(h.p[i - 1], h.p[i - 1], h.p[i]);
   if (!d_jtab) d_jtab = (JumpTable *)dev_alloc(sizeof(JumpTable));
   if (!d_jtab) return false;
   HIP_OK(hipMemcpyAsync(d_jtab, &h, sizeof(h), hipMemcpyHostToDevice, L.stream));
   HIP_OK(hipStreamSynchronize(L.stream));
   jtab_pr = rlx_pr;
   return true;
}

// device memory for the parallel path, or nullptr (the per-stream walk is
// then used) if the device cannot provide it
char *try_alloc(size_t bytes)
{
   void *p = nullptr;
   if (hipMalloc(&p, bytes) != hipSuccess)
   {
      (void)hipGetLastError();
      return nullptr;
   }
   return (char *)p;
}

// one device allocation cut into aligned pieces
struct Carve
{
   size_t total = 0;
   char *base = nullptr;
   size_t take(size_t bytes)
   {
      const size_t o = total;
      total += (bytes + 255) / 256 * 256;
      return o;
   }
   template <typename T> T *at(size_t o) { return (T *)(base + o); }
};

int read_flag(int *bad)
{
   Lib &L = lib();
   int hb = 0;
   HIP_OK(hipMemcpyAsync(&hb, bad, sizeof(int), hipMemcpyDeviceToHost, L.stream));
   HIP_OK(hipStreamSynchronize(L.stream));
   return hb;
}

// the next nu numbers of streams 0 .. nstr-1 into U[s nu + k]
void produce(int nstr, double *U, long long nu, int *bad)
{
   Lib &L = lib();
   const int K = 16;
   const long long nchunk = ((nu + 47) / 48 + K - 1) / K;
   const long long nth = (long long)nstr * nchunk * 4;
   k_rlx_tail<<<nstr, 64, 0, L.stream>>>(d_streams, nstr, U, nu);
   k_rlx_produce<<<(unsigned)((nth + PWG - 1) / PWG), PWG, 0, L.stream>>>(d_streams, nstr, (int)nchunk, K, U, nu,
                                                                       d_jtab, bad);
}

// random_ud with the streams jumped: 0 done, 1 the per-stream walk is needed
// (a link that spans more than the chunk window, the groups ran out before
// the last link, or no device memory for the buffers), -1 error
int random_ud_parallel()
{
   Lib &L = lib();
   const int nstr = L.g.nthread, vt2 = L.g.VT / 2, h = L.g.V / 2;
   const long long nlmax = 8LL * vt2;
   // two groups per link and a margin for the rejected draws (about 2e-4 of
   // the unit vectors and 1e-2 of the second rows are drawn again)
   const long long gneed = 2 * nlmax + nlmax / 16 + 64;
   int C = 256;
   while ((long long)C * 1024 < gneed) C *= 2;
   const long long G = (gneed + C - 1) / C * C;
   if (G * 6 >= (1LL << 31)) return 1;
   const int W = 16, nch = (int)(G / C), sst = (int)nlmax + 1;
   const long long nu = 6 * G, ng = (long long)nstr * G;
   if (!ensure_jtab()) return -1;
   Carve cv;
   const size_t oU = cv.take(sizeof(double) * nstr * nu), oacc = cv.take(ng), oA0 = cv.take(sizeof(int) * ng),
                oA1 = cv.take(sizeof(int) * ng), ork = cv.take(sizeof(int) * nstr * (size_t)vt2),
                oNL = cv.take(sizeof(int) * nstr), oEX = cv.take(sizeof(int) * nstr * (size_t)nch * W),
                oCN = cv.take(sizeof(int) * nstr * (size_t)nch * W), oCB = cv.take(sizeof(int) * nstr * (size_t)nch),
                oCE = cv.take(sizeof(int) * nstr * (size_t)nch), oS = cv.take(sizeof(int) * nstr * (size_t)sst),
                ocons = cv.take(sizeof(long long) * nstr), obad = cv.take(sizeof(int));
   cv.base = try_alloc(cv.total);
   if (!cv.base) return 1;
   double *U = cv.at<double>(oU);
   unsigned char *acc = cv.at<unsigned char>(oacc);
   int *A0 = cv.at<int>(oA0), *A1 = cv.at<int>(oA1), *rank = cv.at<int>(ork), *NL = cv.at<int>(oNL);
   int *EX = cv.at<int>(oEX), *CN = cv.at<int>(oCN), *CB = cv.at<int>(oCB), *CE = cv.at<int>(oCE);
   int *S = cv.at<int>(oS), *bad = cv.at<int>(obad);
   long long *cons = cv.at<long long>(ocons);
   HIP_OK(hipMemsetAsync(bad, 0, sizeof(int), L.stream));
   k_rlx_rank<0><<<nstr, PWG, 0, L.stream>>>(L.d_x0, vt2, h, L.g.L[0], L.bc, 0, rank, NL);
   produce(nstr, U, nu, bad);
   const unsigned gg = (unsigned)((ng + PWG - 1) / PWG);
   k_rlx_unit6<<<gg, PWG, 0, L.stream>>>(U, ng, acc);
   k_rlx_links<<<gg, PWG, 0, L.stream>>>(U, acc, nstr, (int)G, A0, A1);
   const long long nw = (long long)nstr * nch * W;
   k_chain_walk<<<(unsigned)((nw + PWG - 1) / PWG), PWG, 0, L.stream>>>(A1, nstr, (int)G, C, W, EX, CN);
   k_chain_scan<<<(nstr + 63) / 64, 64, 0, L.stream>>>(EX, CN, nstr, nch, W, NL, CB, CE, bad);
   if (read_flag(bad))
   {
      dev_free(cv.base);
      return 1;
   }
   const long long nc = (long long)nstr * nch;
   k_chain_fill<<<(unsigned)((nc + PWG - 1) / PWG), PWG, 0, L.stream>>>(A1, nstr, (int)G, C, CB, CE, NL, S, sst);
   const long long np = (long long)nstr * vt2;
   k_rlx_ud<<<(unsigned)((np + PWG - 1) / PWG), PWG, 0, L.stream>>>(U, A0, A1, S, sst, rank, nstr, vt2, (int)G,
                                                                    L.d_ud, L.d_x0, h, L.g.L[0], L.bc);
   k_ud_cons<<<(nstr + 63) / 64, 64, 0, L.stream>>>(S, sst, NL, nstr, cons);
   k_rlx_finish<<<(nstr + 63) / 64, 64, 0, L.stream>>>(d_streams, nstr, cons, nu, d_jtab, bad);
   const int hb = read_flag(bad);
   dev_free(cv.base);
   if (hb)
   {
      fail("random_ud [uflds.c]", "the random number streams could not be advanced");
      return -1;
   }
   return 0;
}

// random_sd's fill with the streams jumped (0 done, 1 per-stream walk, -1 error)
template <typename C> int random_field_parallel(int nstr, int vol, C *dev, int off, double sigma)
{
   Lib &L = lib();
   const long long nu = 24LL * vol;
   if (!ensure_jtab()) return -1;
   Carve cv;
   const size_t oU = cv.take(sizeof(double) * nstr * nu), ocons = cv.take(sizeof(long long) * nstr),
                obad = cv.take(sizeof(int));
   cv.base = try_alloc(cv.total);
   if (!cv.base) return 1;
   double *U = cv.at<double>(oU);
   long long *cons = cv.at<long long>(ocons);
   int *bad = cv.at<int>(obad);
   HIP_OK(hipMemsetAsync(bad, 0, sizeof(int), L.stream));
   produce(nstr, U, nu, bad);
   if (read_flag(bad))
   {
      dev_free(cv.base);
      return 1;
   }
   const long long np = (long long)nstr * vol;
   k_rlx_sf<C><<<(unsigned)((np + PWG - 1) / PWG), PWG, 0, L.stream>>>(U, nu, nstr, vol, dev, off, sigma);
   k_fill_cons<<<(nstr + 63) / 64, 64, 0, L.stream>>>(cons, nstr, nu);
   k_rlx_finish<<<(nstr + 63) / 64, 64, 0, L.stream>>>(d_streams, nstr, cons, nu, d_jtab, bad);
   const int hb = read_flag(bad);
   dev_free(cv.base);
   if (hb)
   {
      fail("random_sd [sflds.c]", "the random number streams could not be advanced");
      return -1;
   }
   return 0;
}

// the random source with the streams jumped (0 done, 1 per-stream walk, -1 error)
int random_source_parallel(double2 *dev, int t0)
{
   Lib &L = lib();
   const int nstr = L.g.nthread, vt = L.g.VT;
   const long long nu = 24LL * vt;
   if (!ensure_jtab()) return -1;
   Carve cv;
   const size_t oU = cv.take(sizeof(double) * nstr * nu), ork = cv.take(sizeof(int) * (size_t)L.g.V),
                oNL = cv.take(sizeof(int) * nstr), ocons = cv.take(sizeof(long long) * nstr),
                obad = cv.take(sizeof(int));
   cv.base = try_alloc(cv.total);
   if (!cv.base) return 1;
   double *U = cv.at<double>(oU);
   int *rank = cv.at<int>(ork), *NL = cv.at<int>(oNL), *bad = cv.at<int>(obad);
   long long *cons = cv.at<long long>(ocons);
   HIP_OK(hipMemsetAsync(bad, 0, sizeof(int), L.stream));
   k_rlx_rank<1><<<nstr, PWG, 0, L.stream>>>(L.d_x0, vt, 0, L.g.L[0], L.bc, t0, rank, NL);
   produce(nstr, U, nu, bad);
   if (read_flag(bad))
   {
      dev_free(cv.base);
      return 1;
   }
   k_rlx_source<<<(L.g.V + PWG - 1) / PWG, PWG, 0, L.stream>>>(U, nu, rank, L.g.V, vt, dev, L.d_x0, t0);
   k_src_cons<<<(nstr + 63) / 64, 64, 0, L.stream>>>(NL, nstr, cons);
   k_rlx_finish<<<(nstr + 63) / 64, 64, 0, L.stream>>>(d_streams, nstr, cons, nu, d_jtab, bad);
   const int hb = read_flag(bad);
   dev_free(cv.base);
   if (hb)
   {
      fail("oqh_random_source", "the random number streams could not be advanced");
      return -1;
   }
   return 0;
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
   dev_free(d_jtab);
   d_jtab = nullptr;
   jtab_pr = 0;
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
   rlx_pr = h[0].pr;
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
   const int rc = rng_serial() ? 1 : random_ud_parallel();
   if (rc < 0) return;
   if (rc == 1)
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
   const int rc = rng_serial() ? 1 : random_field_parallel<C>(nthr, vol, (C *)r.dev, r.off, sigma);
   if (rc < 0) return;
   if (rc == 1)
      k_random_sf<C><<<(nthr + RWG - 1) / RWG, RWG, 0, L.stream>>>(d_streams, 0, nthr, (C *)r.dev, r.off, vol,
                                                                   sigma);
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
   const int rc = rng_serial() ? 1 : random_source_parallel((double2 *)r.dev, t0);
   if (rc < 0) return;
   if (rc == 1)
      k_random_source<<<(nthr + RWG - 1) / RWG, RWG, 0, L.stream>>>(d_streams, nthr, (double2 *)r.dev, L.d_x0,
                                                                   L.g.VT, t0);
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
