// swterm.hip -- SW (clover) term and its inverse on gfx950.
//
// Replaces sw_term/sw_term.c:225-479 (pauli_term, set_swd, sw_term),
// tcharge/ftensor.c:139-260 (field tensor) and sw_term/pauli_dble.c:1443-1758
// (inv_pauli_dble).  One work-item per point computes the clover
//   F_{mu nu}(x) = (Q_{mu nu}(x) - Q_{mu nu}(x)^dag) / 8
// from the four plaquette leaves at x (gathered, instead of the reference's
// scatter of each plaquette to its four corners), forms the two Hermitian
// 6x6 blocks
//   A_(upper/lower) = c(x0) + c2 * sum_k sigma_k (x) i (B_k -/+ E_k),
//   B = (F23, F31, F12), E = (F01, F02, F03), c2 = -csw/2,
// and stores them in the packed pauli_dble format (6 real diagonal entries,
// then the 15 upper off-diagonal entries row-major).  Points on the boundary
// time slices carry the unit matrix (sw_term.c:293-297); c(x0) = 4+m0, with
// cF-1 added next to the boundaries (sw_term.c:302-330).
//
// The inverse uses Householder QR of (A + i mu) with the reference's safety
// test: the inversion is unsafe if some |R_kk|^2 < 1e-4 |M|_F^2 or if
// |M|_F^2 |M^-1|_F^2 > 1e4 (pauli_dble.c:55-58, 1754-1755).  Only the
// Hermitian part of the inverse is stored.  Each work-item keeps its matrix in
// a private slice of LDS.
#include <cstdio>

#include "oqh_internal.h"

namespace oqh {

__device__ inline void m3_load(const double2 *u, int ox, int l, double2 *m)
{
   const double2 *p = u + (size_t)(ox >> 6) * (72 * SB) + (ox & (SB - 1)) + (size_t)(9 * l) * SB;
#pragma unroll
   for (int k = 0; k < 9; k++) m[k] = p[k * SB];
}

// U(y, mu) for any point y (odd: own slot 2mu; even: slot 2mu+1 of y+mu)
__device__ inline void get_link(const double2 *u, const int *nbr, int V, int y, int mu, double2 *m)
{
   int h = V / 2;
   if (y >= h)
      m3_load(u, y - h, 2 * mu, m);
   else
      m3_load(u, (nbr[(size_t)(2 * mu) * V + y] & 0x7fffffff) - h, 2 * mu + 1, m);
}

__device__ inline void m3_mul(const double2 *a, const double2 *b, double2 *c)
{
   double2 t[9];
#pragma unroll
   for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++)
         t[3 * i + j] = cadd(cadd(cmul(a[3 * i], b[j]), cmul(a[3 * i + 1], b[3 + j])), cmul(a[3 * i + 2], b[6 + j]));
#pragma unroll
   for (int k = 0; k < 9; k++) c[k] = t[k];
}

// c = a * b^dag
__device__ inline void m3_mul_dag(const double2 *a, const double2 *b, double2 *c)
{
   double2 t[9];
#pragma unroll
   for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++)
         t[3 * i + j] = cadd(cadd(cmulc(b[3 * j], a[3 * i]), cmulc(b[3 * j + 1], a[3 * i + 1])),
                             cmulc(b[3 * j + 2], a[3 * i + 2]));
#pragma unroll
   for (int k = 0; k < 9; k++) c[k] = t[k];
}

// c = a^dag * b
__device__ inline void m3_dag_mul(const double2 *a, const double2 *b, double2 *c)
{
   double2 t[9];
#pragma unroll
   for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++)
         t[3 * i + j] = cadd(cadd(cmulc(a[i], b[j]), cmulc(a[3 + i], b[3 + j])), cmulc(a[6 + i], b[6 + j]));
#pragma unroll
   for (int k = 0; k < 9; k++) c[k] = t[k];
}

__device__ inline int nb(const int *nbr, int V, int y, int mu, int dir)
{
   return nbr[(size_t)(2 * mu + dir) * V + y] & 0x7fffffff;
}

// clover sum Q_{mu nu}(x) of the four leaves
__device__ void clover(const double2 *u, const int *nbr, int V, int x, int mu, int nu, double2 *q)
{
   double2 a[9], b[9], c[9], d[9];
   int xpm = nb(nbr, V, x, mu, 0), xpn = nb(nbr, V, x, nu, 0);
   int xmm = nb(nbr, V, x, mu, 1), xmn = nb(nbr, V, x, nu, 1);
   int xpn_mm = nb(nbr, V, xpn, mu, 1);   // x+nu-mu
   int xmm_mn = nb(nbr, V, xmm, nu, 1);   // x-mu-nu
   int xmn_pm = nb(nbr, V, xmn, mu, 0);   // x-nu+mu

   // L1 = U(x,mu) U(x+mu,nu) U(x+nu,mu)^dag U(x,nu)^dag
   get_link(u, nbr, V, x, mu, a);
   get_link(u, nbr, V, xpm, nu, b);
   m3_mul(a, b, c);
   get_link(u, nbr, V, xpn, mu, a);
   m3_mul_dag(c, a, c);
   get_link(u, nbr, V, x, nu, a);
   m3_mul_dag(c, a, q);
   // L2 = U(x,nu) U(x+nu-mu,mu)^dag U(x-mu,nu)^dag U(x-mu,mu)
   get_link(u, nbr, V, x, nu, a);
   get_link(u, nbr, V, xpn_mm, mu, b);
   m3_mul_dag(a, b, c);
   get_link(u, nbr, V, xmm, nu, a);
   m3_mul_dag(c, a, c);
   get_link(u, nbr, V, xmm, mu, a);
   m3_mul(c, a, d);
#pragma unroll
   for (int k = 0; k < 9; k++) q[k] = cadd(q[k], d[k]);
   // L3 = U(x-mu,mu)^dag U(x-mu-nu,nu)^dag U(x-mu-nu,mu) U(x-nu,nu)
   get_link(u, nbr, V, xmm_mn, nu, a);
   get_link(u, nbr, V, xmm, mu, b);
   m3_mul(a, b, c);                       // U(x-mu-nu,nu) U(x-mu,mu)
   get_link(u, nbr, V, xmm_mn, mu, a);
   m3_dag_mul(c, a, c);                   // (..)^dag U(x-mu-nu,mu)
   get_link(u, nbr, V, xmn, nu, a);
   m3_mul(c, a, d);
#pragma unroll
   for (int k = 0; k < 9; k++) q[k] = cadd(q[k], d[k]);
   // L4 = U(x-nu,nu)^dag U(x-nu,mu) U(x-nu+mu,nu) U(x,mu)^dag
   get_link(u, nbr, V, xmn, nu, a);
   get_link(u, nbr, V, xmn, mu, b);
   m3_dag_mul(a, b, c);
   get_link(u, nbr, V, xmn_pm, nu, a);
   m3_mul(c, a, c);
   get_link(u, nbr, V, x, mu, a);
   m3_mul_dag(c, a, d);
#pragma unroll
   for (int k = 0; k < 9; k++) q[k] = cadd(q[k], d[k]);
}

struct SwArgs
{
   const double2 *u;
   double2 *sw;
   double2 *sws;  // the same term as three Hermitian 3x3 matrices per half (or null)
   const int *nbr;
   const int *x0;
   int V, N0, bc;
   double c1, c2, c3a, c3b;
};

__global__ __launch_bounds__(128) void k_sw_term(SwArgs a)
{
   int x = blockIdx.x * blockDim.x + threadIdx.x;
   if (x >= a.V) return;
   double2 *out = a.sw + (size_t)(x >> 6) * (36 * SB) + (x & (SB - 1));
   int t = a.x0[x];
   double2 *so = a.sws ? a.sws + (size_t)(x >> 6) * (27 * SB) + (x & (SB - 1)) : nullptr;
   if ((t == 0 && a.bc != 3) || (t == a.N0 - 1 && a.bc == 0))
   {
#pragma unroll
      for (int k = 0; k < 36; k++)
         out[k * SB] = (k < 3 || (k >= 18 && k < 21)) ? make_double2(1.0, 1.0) : make_double2(0.0, 0.0);
      if (so)
         for (int k = 0; k < 27; k++) so[k * SB] = make_double2(0.0, 0.0);
      return;
   }
   double cx = a.c1;
   if (t == 1 && a.bc != 3) cx = a.c3a;
   else if (t == a.N0 - 2 && a.bc == 0) cx = a.c3b;

   // i*F for the six planes (01),(02),(03),(23),(31),(12) scaled by c2
   const int pm[6] = {0, 0, 0, 2, 3, 1}, pn[6] = {1, 2, 3, 3, 1, 2};
   double2 M[6][9];
   for (int n = 0; n < 6; n++)
   {
      double2 q[9];
      clover(a.u, a.nbr, a.V, x, pm[n], pn[n], q);
#pragma unroll
      for (int i = 0; i < 3; i++)
#pragma unroll
         for (int j = 0; j < 3; j++)
         {
            // F = (q - q^dag)/8 ; M = i c2 F
            double2 f = cscale(csub(q[3 * i + j], make_double2(q[3 * j + i].x, -q[3 * j + i].y)), 0.125);
            M[n][3 * i + j] = make_double2(-a.c2 * f.y, a.c2 * f.x);
         }
   }
   double dgs[18];
   for (int half = 0; half < 2; half++)
   {
      double sg = half ? 1.0 : -1.0;
      double2 S[3][9];  // M_k = i c2 (B_k + sg E_k)
#pragma unroll
      for (int k = 0; k < 3; k++)
#pragma unroll
         for (int e = 0; e < 9; e++) S[k][e] = cadd(M[3 + k][e], cscale(M[k][e], sg));
      // H = [[S3, S1 - i S2], [S1 + i S2, -S3]]
      double2 H[36];
#pragma unroll
      for (int i = 0; i < 3; i++)
#pragma unroll
         for (int j = 0; j < 3; j++)
         {
            double2 s1 = S[0][3 * i + j], s2 = S[1][3 * i + j], s3 = S[2][3 * i + j];
            H[6 * i + j] = s3;
            H[6 * (i + 3) + j + 3] = make_double2(-s3.x, -s3.y);
            H[6 * i + j + 3] = csubi(s1, s2);
            H[6 * (i + 3) + j] = caddi(s1, s2);
         }
      if (so)
      {
         // S-form: units 0-8 the real diagonals of S1, S2, S3 of both halves
         // (18 reals, written after the loop), units 9 + 9 half + 3 k + q the
         // entries (0,1), (0,2), (1,2) of S_{k+1}
#pragma unroll
         for (int k = 0; k < 3; k++)
#pragma unroll
            for (int i = 0; i < 3; i++) dgs[9 * half + 3 * k + i] = S[k][4 * i].x;
#pragma unroll
         for (int k = 0; k < 3; k++)
         {
            so[(9 + 9 * half + 3 * k + 0) * SB] = S[k][1];
            so[(9 + 9 * half + 3 * k + 1) * SB] = S[k][2];
            so[(9 + 9 * half + 3 * k + 2) * SB] = S[k][5];
         }
      }
      double2 *o = out + (size_t)(18 * half) * SB;
      o[0] = make_double2(H[0].x + cx, H[7].x + cx);
      o[SB] = make_double2(H[14].x + cx, H[21].x + cx);
      o[2 * SB] = make_double2(H[28].x + cx, H[35].x + cx);
      int p = 3;
#pragma unroll
      for (int i = 0; i < 6; i++)
#pragma unroll
         for (int j = i + 1; j < 6; j++) o[(p++) * SB] = H[6 * i + j];
   }
   if (so)
#pragma unroll
      for (int r = 0; r < 9; r++) so[r * SB] = make_double2(dgs[2 * r], dgs[2 * r + 1]);
}

// ---------------------------------------------------------------------------
// inversion of the Pauli blocks
// ---------------------------------------------------------------------------

constexpr int INV_WG = 32;  // 32 x 72 x 16 B = 36.9 KB of LDS per work-group

struct InvArgs
{
   const double2 *in;
   double2 *out;
   int p0, np;   // points
   double mu;
   int *ifail;
};

__global__ __launch_bounds__(INV_WG) void k_inv_pauli(InvArgs a)
{
   __shared__ double2 lds[INV_WG][2][36];
   int k0 = blockIdx.x * INV_WG + threadIdx.x;
   if (k0 >= 2 * a.np) return;
   int x = a.p0 + (k0 >> 1), half = k0 & 1;
   double2 *A = lds[threadIdx.x][0];
   double2 *X = lds[threadIdx.x][1];
   const double2 *m = a.in + (size_t)(x >> 6) * (36 * SB) + (x & (SB - 1)) + (size_t)(18 * half) * SB;
   double mu = half ? -a.mu : a.mu;

   // unpack (A + i mu)
   double2 d0 = m[0], d1 = m[SB], d2 = m[2 * SB];
   double dg[6] = {d0.x, d0.y, d1.x, d1.y, d2.x, d2.y};
   double nrm = 0.0;
   for (int i = 0; i < 6; i++)
   {
      A[7 * i] = make_double2(dg[i], mu);
      nrm += dg[i] * dg[i] + mu * mu;
   }
   int p = 3;
   for (int i = 0; i < 6; i++)
      for (int j = i + 1; j < 6; j++)
      {
         double2 v = m[(p++) * SB];
         A[6 * i + j] = v;
         A[6 * j + i] = make_double2(v.x, -v.y);
         nrm += 2.0 * (v.x * v.x + v.y * v.y);
      }
   double eps = 1.0e-4 * nrm;
   int fail = 0;
   double beta[6];
   double2 rd[6];

   // Householder reflections; v stored in column k (rows >= k)
   for (int k = 0; k < 6; k++)
   {
      double n2 = 0.0;
      for (int i = k; i < 6; i++) n2 += A[6 * i + k].x * A[6 * i + k].x + A[6 * i + k].y * A[6 * i + k].y;
      if (n2 < eps) fail = 1;
      beta[k] = 0.0;
      if (k == 5 || n2 == 0.0)
      {
         rd[k] = A[6 * k + k];
         continue;
      }
      double al = sqrt(n2);
      double2 a0 = A[6 * k + k];
      double aa = sqrt(a0.x * a0.x + a0.y * a0.y);
      double2 ph = (aa > 0.0) ? make_double2(a0.x / aa, a0.y / aa) : make_double2(1.0, 0.0);
      A[6 * k + k] = cadd(a0, cscale(ph, al));
      beta[k] = 1.0 / (al * (al + aa));
      rd[k] = cscale(ph, -al);
      for (int j = k + 1; j < 6; j++)
      {
         double2 s = make_double2(0.0, 0.0);
         for (int i = k; i < 6; i++) s = cadd(s, cmulc(A[6 * i + k], A[6 * i + j]));
         s = cscale(s, beta[k]);
         for (int i = k; i < 6; i++) A[6 * i + j] = csub(A[6 * i + j], cmul(s, A[6 * i + k]));
      }
   }
   // columns of M^-1 = R^-1 Q^dag e_j
   double ninv = 0.0;
   for (int j = 0; j < 6; j++)
   {
      double2 b[6];
      for (int i = 0; i < 6; i++) b[i] = make_double2(i == j ? 1.0 : 0.0, 0.0);
      for (int k = 0; k < 5; k++)
      {
         if (beta[k] == 0.0) continue;
         double2 s = make_double2(0.0, 0.0);
         for (int i = k; i < 6; i++) s = cadd(s, cmulc(A[6 * i + k], b[i]));
         s = cscale(s, beta[k]);
         for (int i = k; i < 6; i++) b[i] = csub(b[i], cmul(s, A[6 * i + k]));
      }
      for (int i = 5; i >= 0; i--)
      {
         double2 s = b[i];
         for (int l = i + 1; l < 6; l++) s = csub(s, cmul(A[6 * i + l], X[6 * l + j]));
         double r2 = rd[i].x * rd[i].x + rd[i].y * rd[i].y;
         X[6 * i + j] = (r2 > 0.0) ? cscale(cmulc(rd[i], s), 1.0 / r2) : make_double2(0.0, 0.0);
         ninv += X[6 * i + j].x * X[6 * i + j].x + X[6 * i + j].y * X[6 * i + j].y;
      }
   }
   if (eps * ninv > 1.0) fail = 1;
   if (fail) atomicOr(a.ifail, 1);

   // Hermitian part, packed
   double2 *o = a.out + (size_t)(x >> 6) * (36 * SB) + (x & (SB - 1)) + (size_t)(18 * half) * SB;
   o[0] = make_double2(X[0].x, X[7].x);
   o[SB] = make_double2(X[14].x, X[21].x);
   o[2 * SB] = make_double2(X[28].x, X[35].x);
   p = 3;
   for (int i = 0; i < 6; i++)
      for (int j = i + 1; j < 6; j++)
      {
         double2 zij = X[6 * i + j], zji = X[6 * j + i];
         o[(p++) * SB] = make_double2(0.5 * (zij.x + zji.x), 0.5 * (zij.y - zji.y));
      }
}

static int *d_ifail = nullptr;

int invert_pauli(const double2 *in, double2 *out, int p0, int np, double mu)
{
   Lib &L = lib();
   if (!d_ifail) d_ifail = (int *)dev_alloc(sizeof(int));
   HIP_OK(hipMemsetAsync(d_ifail, 0, sizeof(int), L.stream));
   InvArgs a{in, out, p0, np, mu, d_ifail};
   unsigned nbk = (unsigned)((2 * (size_t)np + INV_WG - 1) / INV_WG);
   if (nbk) k_inv_pauli<<<nbk, INV_WG, 0, L.stream>>>(a);
   int h = 0;
   HIP_OK(hipMemcpyAsync(&h, d_ifail, sizeof(int), hipMemcpyDeviceToHost, L.stream));
   HIP_OK(hipStreamSynchronize(L.stream));
   return h;
}

static void compute_swd()
{
   Lib &L = lib();
   SwArgs a;
   a.u = L.d_ud;
   a.sw = L.d_swd;
   if (!L.d_sws) L.d_sws = (double2 *)dev_alloc(units_for(L.g.V, 27) * sizeof(double2));
   a.sws = L.d_sws;
   a.nbr = L.d_nbr;
   a.x0 = L.d_x0;
   a.V = L.g.V;
   a.N0 = L.g.L[0];
   a.bc = L.bc;
   a.c1 = 4.0 + L.m0;
   a.c2 = -0.5 * L.csw;
   a.c3a = a.c1 + L.cF[0] - 1.0;
   a.c3b = a.c1 + L.cF[1] - 1.0;
   k_sw_term<<<(unsigned)((L.g.V + 127) / 128), 128, 0, L.stream>>>(a);
   L.sws_valid = L.d_sws != nullptr;
   L.sws32_valid = false;
   L.swd_valid = true;
   L.swd_inv[0] = L.swd_inv[1] = false;
   L.sw_valid = false;
   L.swd_epoch++;
}

// make the double-precision SW field current, with the even/odd parts
// inverted as requested; returns nonzero if an inversion was unsafe
int ensure_swd_ret(int want_e, int want_o)
{
   Lib &L = lib();
   if (!L.ud_set)
   {
      fail("sw_term [sw_term.c]", "gauge field is not set");
      return 1;
   }
   if (!L.swd_valid || (L.swd_inv[0] && !want_e) || (L.swd_inv[1] && !want_o)) compute_swd();
   int ifail = 0, h = L.g.V / 2;
   if (want_e && !L.swd_inv[0])
   {
      ifail |= invert_pauli(L.d_swd, L.d_swd, 0, h, 0.0);
      L.swd_inv[0] = true;
      L.sw_valid = false;
      L.swd_epoch++;
   }
   if (want_o && !L.swd_inv[1])
   {
      ifail |= invert_pauli(L.d_swd, L.d_swd, h, h, 0.0);
      L.swd_inv[1] = true;
      L.sw_valid = false;
      L.swd_epoch++;
   }
   return ifail;
}

void ensure_swd(int want_e, int want_o) { (void)ensure_swd_ret(want_e, want_o); }

void ensure_sw()
{
   Lib &L = lib();
   if (!L.swd_valid) ensure_swd(0, 0);
   if (L.sw_valid) return;
   convert_f64_to_f32(L.d_swd, L.d_sw, units_for(L.g.V, 36), L.stream);
   if (L.sws_valid && !L.sws32_valid)
   {
      if (!L.d_sws32) L.d_sws32 = (float2 *)dev_alloc(units_for(L.g.V, 27) * sizeof(float2));
      if (L.d_sws32)
      {
         convert_f64_to_f32(L.d_sws, L.d_sws32, units_for(L.g.V, 27), L.stream);
         L.sws32_valid = true;
      }
   }
   L.sw_valid = true;
}

void ensure_u()
{
   Lib &L = lib();
   if (L.u_valid) return;
   convert_f64_to_f32(L.d_ud, L.d_u, units_for(L.g.V / 2, 72), L.stream);
   L.u_valid = true;
   L.u_epoch++;
}

}  // namespace oqh

using namespace oqh;

extern "C" int sw_term(ptset_t set)
{
   if (!require_geometry("sw_term [sw_term.c]")) return 1;
   int ie = (set == ALL_PTS || set == EVEN_PTS), io = (set == ALL_PTS || set == ODD_PTS);
   return ensure_swd_ret(ie, io);
}

extern "C" void assign_swd2sw(void)
{
   Lib &L = lib();
   if (!require_geometry("assign_swd2sw [swflds.c]")) return;
   L.sw_valid = false;
   ensure_sw();
}

extern "C" void assign_ud2u(void)
{
   Lib &L = lib();
   if (!require_geometry("assign_ud2u [uflds.c]")) return;
   L.u_valid = false;
   ensure_u();
}

// swflds.c:79-146: the SW fields.  The fields live on the device; swdfld()
// and swfld() return host copies (2V matrices in the reference's layout)
// taken at the time of the call: the current SW term as last computed
// (uninverted or with the inverted parts sw_term() left), or unit matrices
// when no SW term has been computed.  Changes made through the returned
// pointers do not reach the device (oqh_import_swd does).
static std::vector<pauli_dble> h_swd;
static std::vector<pauli> h_sw;

extern "C" pauli_dble *swdfld(void)
{
   Lib &L = lib();
   if (!require_geometry("swdfld [swflds.c]")) return nullptr;
   h_swd.assign(2 * (size_t)L.g.V, pauli_dble{});
   if (L.swd_valid)
      oqh_export_swd(h_swd.data());
   else
      for (auto &m : h_swd)
         for (int k = 0; k < 6; k++) m.u[k] = 1.0;
   return h_swd.data();
}

extern "C" pauli *swfld(void)
{
   Lib &L = lib();
   if (!require_geometry("swfld [swflds.c]")) return nullptr;
   h_sw.assign(2 * (size_t)L.g.V, pauli{});
   if (L.swd_valid)
   {
      ensure_sw();
      const size_t n = (size_t)L.g.V;
      float2 *aos = (float2 *)tmp_buffer(n * sizeof(pauli) * 2);
      if (!aos) return nullptr;
      blocks_to_aos_f32(L.d_sw, aos, L.g.V, 36, L.stream);
      HIP_OK(hipMemcpyAsync(h_sw.data(), aos, n * 2 * sizeof(pauli), hipMemcpyDeviceToHost, L.stream));
      HIP_OK(hipStreamSynchronize(L.stream));
   }
   else
      for (auto &m : h_sw)
         for (int k = 0; k < 6; k++) m.u[k] = 1.0f;
   return h_sw.data();
}
