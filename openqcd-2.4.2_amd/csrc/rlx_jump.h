// rlx_jump.h -- skip-ahead of the 48-bit subtract-with-borrow lanes of ranlxd.
//
// One lane of ranlxd (random/ranlxd.c) is the recursion x_n = x_{n-5} -
// x_{n-12} - c_{n-1} mod b, b = 2^48, with the borrow c_n.  It is the linear
// congruential generator Z_n = a Z_{n-1} mod m, m = b^12 - b^5 + 1, a = b^-1
// mod m (A. Sibidanov, "A revision of the subtract-with-borrow random number
// generators", Comput. Phys. Commun. 221 (2017) 299), where
//
//    Z_n = sum_{k<12} x_{n-11+k} b^k - sum_{k<5} x_{n-4+k} b^k + c_n
//
// and the words come back as base-b digits: x_n = floor(b Z_n / m), Z_{n-1} =
// b Z_n mod m.  (Proof: the recursion gives Z_{n-1} = b Z_n - x_n m.)  A jump
// by N steps is one product with a^N mod m.  Residues are 19 little-endian
// 32-bit limbs (the top limb is zero for reduced values).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace oqh {
namespace rlx {

constexpr int NR = 19;  // limbs of a residue (576 bits + one spare limb)

// m = 2^576 - 2^240 + 1
__host__ __device__ inline void modulus(uint32_t *m)
{
   for (int i = 0; i < NR; i++) m[i] = 0;
   m[0] = 1;
   m[7] = 0xffff0000u;
   for (int i = 8; i < 18; i++) m[i] = 0xffffffffu;
}

__host__ __device__ inline bool geq(const uint32_t *x, const uint32_t *y)
{
   for (int i = NR - 1; i >= 0; i--)
      if (x[i] != y[i]) return x[i] > y[i];
   return true;
}

// x -= y (x >= y)
__host__ __device__ inline void sub(uint32_t *x, const uint32_t *y)
{
   int64_t bw = 0;
   for (int i = 0; i < NR; i++)
   {
      const int64_t d = (int64_t)x[i] - (int64_t)y[i] + bw;
      x[i] = (uint32_t)d;
      bw = (d < 0) ? -1 : 0;
   }
}

// x += v 2^bit (v < 2^48, bit % 32 in {0, 16}) over n limbs
__host__ __device__ inline void add_at(uint32_t *x, int n, int bit, uint64_t v)
{
   const int i0 = bit >> 5;
   const uint64_t w = v << (bit & 31);
   uint64_t c = 0;
   for (int i = i0; i < n; i++)
   {
      const uint64_t add = (i == i0) ? (w & 0xffffffffu) : (i == i0 + 1) ? (w >> 32) : 0;
      const uint64_t p = (uint64_t)x[i] + add + c;
      x[i] = (uint32_t)p;
      c = p >> 32;
      if (i > i0 && c == 0) break;
   }
}

// x -= v 2^bit (same conditions, x >= v 2^bit)
__host__ __device__ inline void sub_at(uint32_t *x, int n, int bit, uint64_t v)
{
   const int i0 = bit >> 5;
   const uint64_t w = v << (bit & 31);
   int64_t bw = 0;
   for (int i = i0; i < n; i++)
   {
      const int64_t s = (i == i0) ? (int64_t)(w & 0xffffffffu) : (i == i0 + 1) ? (int64_t)(w >> 32) : 0;
      const int64_t d = (int64_t)x[i] - s + bw;
      x[i] = (uint32_t)d;
      bw = (d < 0) ? -1 : 0;
      if (i > i0 && bw == 0) break;
   }
}

// t (38 limbs) -> L + H 2^240 - H with L = t mod 2^576, H = t >> 576, which
// is congruent modulo m since 2^576 = 2^240 - 1 mod m
__host__ __device__ inline void fold(uint32_t *t)
{
   uint32_t h[20];
   for (int i = 0; i < 20; i++)
   {
      h[i] = t[18 + i];
      t[18 + i] = 0;
   }
   uint64_t c = 0;
   for (int i = 0; i < 38; i++)
   {
      const int k = i - 7;  // 240 = 7 limbs + 16 bits
      uint32_t s = 0;
      if (k >= 0 && k < 20) s |= h[k] << 16;
      if (k >= 1 && k <= 20) s |= h[k - 1] >> 16;
      const uint64_t p = (uint64_t)t[i] + s + c;
      t[i] = (uint32_t)p;
      c = p >> 32;
   }
   int64_t bw = 0;
   for (int i = 0; i < 38; i++)
   {
      const int64_t d = (int64_t)t[i] - (int64_t)(i < 20 ? h[i] : 0u) + bw;
      t[i] = (uint32_t)d;
      bw = (d < 0) ? -1 : 0;
   }
}

// r = x y mod m (x, y reduced; r may alias x or y)
__host__ __device__ inline void mulmod(const uint32_t *x, const uint32_t *y, uint32_t *r)
{
   uint32_t t[38];
   for (int i = 0; i < 38; i++) t[i] = 0;
   for (int i = 0; i < 18; i++)
   {
      uint64_t c = 0;
      for (int j = 0; j < 18; j++)
      {
         const uint64_t p = (uint64_t)x[i] * y[j] + t[i + j] + c;
         t[i + j] = (uint32_t)p;
         c = p >> 32;
      }
      t[i + 18] = (uint32_t)c;
   }
   fold(t);  // < 2^576 + 2^816
   fold(t);  // < 2^577
   uint32_t m[NR];
   modulus(m);
   for (int i = 0; i < NR; i++) r[i] = t[i];
   while (geq(r, m)) sub(r, m);
}

// P - Q of the words x[0..11] (chronological, x[11] newest): 19 limbs
__host__ __device__ inline void words_pq(const uint64_t *x, uint32_t *z)
{
   for (int i = 0; i < NR; i++) z[i] = 0;
   for (int k = 0; k < 12; k++) add_at(z, NR, 48 * k, x[k]);
   for (int k = 0; k < 5; k++) sub_at(z, NR, 48 * k, x[7 + k]);
}

// reduced Z of the state (words x[0..11] chronological, borrow c)
__host__ __device__ inline void state_to_z(const uint64_t *x, int c, uint32_t *z)
{
   words_pq(x, z);
   add_at(z, NR, 0, (uint64_t)c);
   uint32_t m[NR];
   modulus(m);
   while (geq(z, m)) sub(z, m);
}

// words x[0..11] (chronological) and borrow of the state with residue z;
// returns false if the borrow is not 0 or 1 (a state off the generator's
// orbit)
__host__ __device__ inline bool z_to_state(const uint32_t *z, uint64_t *x, int *c)
{
   uint32_t m[NR], y[NR];
   modulus(m);
   for (int i = 0; i < NR; i++) y[i] = z[i];
   for (int k = 11; k >= 0; k--)
   {
      // q = floor(b y / m) and y <- b y mod m: b y = q 2^576 + low, 2^576 =
      // m + 2^240 - 1, so b y - q m = low + q 2^240 - q
      const uint64_t q = ((uint64_t)y[17] << 16) | (y[16] >> 16);
      uint32_t t[NR];
      t[18] = 0;
      for (int i = 17; i >= 0; i--)
      {
         const uint32_t hi = (i >= 1) ? y[i - 1] : 0u, lo = (i >= 2) ? y[i - 2] : 0u;
         t[i] = (hi << 16) | (lo >> 16);  // (y << 48) mod 2^576
      }
      add_at(t, NR, 240, q);
      sub_at(t, NR, 0, q);
      uint64_t qq = q;
      if (geq(t, m))
      {
         sub(t, m);
         qq++;
      }
      x[k] = qq;
      for (int i = 0; i < NR; i++) y[i] = t[i];
   }
   uint32_t d[NR];
   words_pq(x, d);
   // borrow = z - (P - Q), taken modulo m
   uint32_t e[NR];
   for (int i = 0; i < NR; i++) e[i] = z[i];
   if (!geq(e, d))
   {
      uint64_t cc = 0;
      for (int i = 0; i < NR; i++)
      {
         const uint64_t p = (uint64_t)e[i] + m[i] + cc;
         e[i] = (uint32_t)p;
         cc = p >> 32;
      }
   }
   sub(e, d);
   for (int i = 1; i < NR; i++)
      if (e[i]) return false;
   if (e[0] > 1) return false;
   *c = (int)e[0];
   return true;
}

// a = b^-1 mod m = m - (b^11 - b^4)
__host__ __device__ inline void multiplier(uint32_t *a)
{
   modulus(a);
   uint32_t t[NR];
   for (int i = 0; i < NR; i++) t[i] = 0;
   // b^11 - b^4 = 2^528 - 2^192: bits 192..527
   for (int i = 6; i < 16; i++) t[i] = 0xffffffffu;
   t[16] = 0xffffu;
   sub(a, t);
}

}  // namespace rlx
}  // namespace oqh
