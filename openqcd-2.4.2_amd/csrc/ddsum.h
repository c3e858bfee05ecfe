// ddsum.h -- double-double sums on the device: the addition of two
// double-double numbers with the steps of add_qflt (utils/qsum.c:284-310) and
// a work-group reduction built on it.
#pragma once

#include <hip/hip_runtime.h>

namespace oqh {

struct dd
{
   double hi, lo;
};

__device__ __host__ inline dd dd_add(dd a, dd b)
{
   // sum of two double-double numbers (same steps as add_qflt, qsum.c:284)
   double s = a.hi + b.hi;
   double c = a.lo + b.lo;
   double up = s - b.hi;
   double rp = c - b.lo;
   double vp = s - up;
   double sp = c - rp;
   double e1 = (a.hi - up) + (b.hi - vp);
   double d = (a.lo - rp) + (b.lo - sp);
   e1 += c;
   double e = s + e1;
   double f = e1 - (e - s);
   f += d;
   double w = e + f;
   dd r;
   r.hi = w;
   r.lo = f - (w - e);
   return r;
}

__device__ inline dd shfl_xor_dd(dd v, int m)
{
   dd r;
   r.hi = __shfl_xor(v.hi, m);
   r.lo = __shfl_xor(v.lo, m);
   return r;
}

// reduce NV double-double values over the work-group; result in thread 0
template <int NV, int RB = 256> __device__ inline void block_reduce_dd(dd *v)
{
   __shared__ dd sh[RB / 64][NV];
#pragma unroll
   for (int m = 32; m >= 1; m >>= 1)
#pragma unroll
      for (int k = 0; k < NV; k++) v[k] = dd_add(v[k], shfl_xor_dd(v[k], m));
   int w = threadIdx.x >> 6, l = threadIdx.x & 63;
   if (l == 0)
#pragma unroll
      for (int k = 0; k < NV; k++) sh[w][k] = v[k];
   __syncthreads();
   if (threadIdx.x == 0)
   {
      for (int ww = 1; ww < RB / 64; ww++)
#pragma unroll
         for (int k = 0; k < NV; k++) v[k] = dd_add(v[k], sh[ww][k]);
   }
}

}  // namespace oqh
